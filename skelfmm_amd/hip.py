"""ctypes binding of ``libskelfmm_hip.so`` (declared in ``include/skelfmm_hip.h``).

The library is loaded from this package directory. There is no CPU fallback:
if the library or a GPU is missing, every entry point raises ``RuntimeError``.
Device memory comes from PyTorch tensors; launches go to PyTorch's current
HIP stream.
"""

from __future__ import annotations

import ctypes
import os

import torch

__all__ = ["lib", "torch_ops", "check", "stream_handle", "P2PArgs", "P2PLists", "LevelArgs", "PlanDesc", "p2p_blocks", "LIB_PATH",
           "EXPORTED_SYMBOLS"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libskelfmm_hip.so")
TORCH_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libskelfmm_torch.so")

EXPORTED_SYMBOLS = (
    "skf_last_error",
    "skf_version",
    "skf_build_id",
    "skf_fill_zero",
    "skf_gather_charges",
    "skf_scatter_result",
    "skf_p2p",
    "skf_p2p_noalloc",
    "skf_up_level",
    "skf_down_level",
    "skf_plan_create",
    "skf_plan_destroy",
    "skf_apply",
    "skf_p2p_blocks",
    "skf_build_last_error",
    "skf_proxy_matrices",
    "skf_pqr_batched",
    "skf_trsm_batched",
    "skf_id_residual",
    "skf_mlp_param_count",
    "skf_mlp_forward",
    "skf_mlp_backward",
    "skf_adam_step",
    "skf_jit_compile",
    "skf_jit_log",
    "skf_jit_unload",
    "skf_tree_last_error",
    "skf_tree_keys",
    "skf_tree_sort_pairs",
    "skf_tree_child_ranges",
    "skf_tree_balance_marks",
    "skf_tree_box_tables",
    "skf_tree_colleagues",
    "skf_tree_coarse",
    "skf_tree_pairs",
)

_vp = ctypes.c_void_p
_i = ctypes.c_int


class P2PLists(ctypes.Structure):
    _fields_ = [
        ("ngroups", _i), ("g_pt", _vp), ("g_n", _vp), ("g_out", _vp), ("g_box", _vp),
        ("nentries", _i), ("e_pt", _vp), ("e_n", _vp), ("e_q", _vp), ("e_box", _vp), ("e_flags", _vp),
        ("nitems", _i), ("items", _vp), ("e_shift", _vp),
    ]


class P2PArgs(ctypes.Structure):
    _fields_ = [
        ("kernel", _i), ("precision", _i), ("width", _i), ("kappa", ctypes.c_double),
        ("px", _vp), ("py", _vp), ("pz", _vp), ("nx", _vp), ("ny", _vp), ("nz", _vp),
        ("centers", _vp), ("q", _vp), ("u", _vp),
        ("ngroups", _i), ("g_pt", _vp), ("g_n", _vp), ("g_out", _vp), ("g_box", _vp),
        ("nentries", _i), ("e_pt", _vp), ("e_n", _vp), ("e_q", _vp), ("e_box", _vp), ("e_flags", _vp),
        ("nitems", _i), ("items", _vp), ("mfma_min_channels", _i),
        ("one", ctypes.POINTER(P2PLists)), ("sym", ctypes.POINTER(P2PLists)), ("urev", _vp),
        ("urev_len", ctypes.c_longlong), ("sym_unroll", _i), ("user_fn", _vp * 4),
        ("max_distance", ctypes.c_double), ("geo64", _vp), ("geo64_n", ctypes.c_longlong), ("e_shift", _vp),
        ("store_rows", _i),
    ]


class LevelArgs(ctypes.Structure):
    _fields_ = [
        ("nboxes", _i), ("box_off", _vp), ("box_bs", _vp), ("box_rank", _vp), ("box_toff", _vp),
        ("ngather", _i), ("gather_dst", _vp), ("gather_src", _vp),
        ("nitems", _i), ("items", _vp), ("variant", _i),
    ]


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("precision", _i), ("complex_t", _i), ("width", _i), ("nslots", _i), ("npoints", _i),
        ("leaf_slot", _vp), ("leaf_user", _vp), ("T", _vp), ("qbuf", _vp), ("ubuf", _vp),
        ("nup", _i), ("up", ctypes.POINTER(LevelArgs)), ("ndown", _i), ("down", ctypes.POINTER(LevelArgs)),
        ("p2p", P2PArgs),
    ]


_LIB = None


def _ensure_current():
    """Make sure the library file was built from the sources beside it (the hash of csrc/ and
    include/ embedded at build time); a missing or stale library is rebuilt unless
    ``SKELFMM_AMD_NO_REBUILD`` is set, in which case it is an error."""
    from . import _build

    want = _build.source_hash()
    have = _build.built_hash(LIB_PATH)
    if have == want:
        return want
    what = "is not built" if have is None else f"was built from other sources ({have}, sources {want})"
    if os.environ.get("SKELFMM_AMD_NO_REBUILD"):
        raise RuntimeError(f"HIP library {LIB_PATH} {what}; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    import sys

    print(f"skelfmm_amd: HIP library {what}; compiling it for gfx950", file=sys.stderr, flush=True)
    _build.build_library()
    return want


_TORCH_OPS = None


def torch_ops():
    """The registered PyTorch operators ``torch.ops.skelfmm`` (``apply(plan, x, y)``: the whole
    apply of a plan on the current stream), from ``libskelfmm_torch.so``, a thin C++ extension
    over the C ABI built with the HIP library; rebuilt when its embedded source hash differs from
    the sources (unless ``SKELFMM_AMD_NO_REBUILD`` is set: then an error)."""
    global _TORCH_OPS
    if _TORCH_OPS is not None:
        return _TORCH_OPS
    lib()  # the operator library links the HIP library
    from . import _build

    want = _build.source_hash()
    have = _build.built_hash(TORCH_LIB_PATH)
    if have != want:
        if os.environ.get("SKELFMM_AMD_NO_REBUILD"):
            raise RuntimeError(f"PyTorch operator library {TORCH_LIB_PATH} is missing or stale ({have}, sources "
                               f"{want}); run __graft_entry__.build()")
        _build.build_torch_ops()
    torch.ops.load_library(TORCH_LIB_PATH)
    _TORCH_OPS = torch.ops.skelfmm
    return _TORCH_OPS


def build_id():
    """Source hash of the loaded library (equal to that of the sources beside it)."""
    return lib().skf_build_id().decode()


def lib():
    """Load the HIP library, compiling it first when it is missing or older than its sources."""
    global _LIB
    if _LIB is not None:
        return _LIB
    want = _ensure_current()
    L = ctypes.CDLL(LIB_PATH)
    L.skf_build_id.restype = ctypes.c_char_p
    if L.skf_build_id().decode() != want:
        raise RuntimeError(f"HIP library {LIB_PATH}: loaded build id {L.skf_build_id().decode()} != {want}")
    L.skf_last_error.restype = ctypes.c_char_p
    L.skf_version.restype = _i
    L.skf_fill_zero.argtypes = [_vp, ctypes.c_longlong, _i, _vp]
    L.skf_gather_charges.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _i, _vp]
    L.skf_scatter_result.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _i, _vp]
    L.skf_p2p.argtypes = [ctypes.POINTER(P2PArgs), _vp]
    L.skf_p2p_noalloc.argtypes = [ctypes.POINTER(P2PArgs), _vp]
    L.skf_up_level.argtypes = [ctypes.POINTER(LevelArgs), _vp, _i, _vp, _vp, _i, _i, _vp]
    L.skf_down_level.argtypes = [ctypes.POINTER(LevelArgs), _vp, _i, _vp, _vp, _i, _i, _vp]
    L.skf_plan_create.argtypes = [ctypes.POINTER(PlanDesc), ctypes.POINTER(_vp)]
    L.skf_plan_destroy.argtypes = [_vp]
    L.skf_apply.argtypes = [_vp, _vp, _vp, _vp]
    L.skf_p2p_blocks.argtypes = [_i, _i, _i, ctypes.c_double, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp]
    L.skf_build_last_error.restype = ctypes.c_char_p
    L.skf_proxy_matrices.argtypes = [_i, _i, ctypes.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i,
                                     _vp, _vp, _i, _i, _i, _vp]
    L.skf_pqr_batched.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp, _vp, _vp, _i, _i,
                                  _i, _i, _vp]
    L.skf_trsm_batched.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp]
    L.skf_id_residual.argtypes = [_vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp]
    L.skf_mlp_forward.argtypes = [_vp, _vp, _i, _vp, _vp]
    L.skf_mlp_backward.argtypes = [_vp, _vp, _vp, _i, _vp, _vp, _vp]
    L.skf_adam_step.argtypes = [_vp, _vp, _vp, _vp, _i, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, _i, _vp]
    L.skf_jit_compile.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p),
                                  _i, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_longlong)]
    L.skf_jit_log.restype = ctypes.c_char_p
    L.skf_jit_unload.argtypes = [_vp]
    L.skf_tree_last_error.restype = ctypes.c_char_p
    _ll = ctypes.c_longlong
    L.skf_tree_keys.argtypes = [_vp, _i, _i, _vp, ctypes.c_double, _i, _vp, _vp, _vp]
    L.skf_tree_sort_pairs.argtypes = [_vp, ctypes.POINTER(ctypes.c_ulonglong), _vp, _vp, _vp, _vp, _i, _i, _vp]
    L.skf_tree_child_ranges.argtypes = [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.skf_tree_balance_marks.argtypes = [_i, _i, _ll, _vp, _vp, _vp, _vp, _vp]
    L.skf_tree_box_tables.argtypes = [_i, _i, _ll, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.skf_tree_colleagues.argtypes = [_i, _vp, _vp, _vp, _ll, _ll, _vp, _vp]
    L.skf_tree_coarse.argtypes = [_i, _vp, _vp, _vp, _vp, _ll, _ll, _vp, _vp]
    L.skf_tree_pairs.argtypes = [_i, _vp, _vp, _vp, _ll, _ll, _i, _vp, _vp, _vp, _vp, _vp, _vp]
    for name in EXPORTED_SYMBOLS:
        if name not in ("skf_last_error", "skf_version", "skf_build_id", "skf_build_last_error", "skf_jit_log",
                        "skf_tree_last_error"):
            getattr(L, name).restype = _i
    _LIB = L
    return L


def check(rc):
    if rc != 0:
        raise RuntimeError("skelfmm_hip: " + lib().skf_last_error().decode(errors="replace"))


def check_build(rc):
    if rc != 0:
        raise RuntimeError("skelfmm_hip (build): " + lib().skf_build_last_error().decode(errors="replace"))


def check_tree(rc):
    if rc != 0:
        raise RuntimeError("skelfmm_hip (tree): " + lib().skf_tree_last_error().decode(errors="replace"))


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("skelfmm_amd applies run on the GPU (HIP); no GPU is visible to PyTorch")


def stream_handle(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def p2p_blocks(kernel_id, q, x_trg, x_src, param, self_interaction, geom_trg=None, geom_src=None):
    """Dense block evaluation ``(B, m, G) = K(x_trg, x_src) @ q`` on the GPU (operator ``torch_apply``)."""
    require_gpu()
    dev = x_trg.device if x_trg.is_cuda else torch.device("cuda")
    real = torch.float64 if x_trg.dtype == torch.float64 else torch.float32
    squeeze = x_trg.dim() == 2
    if squeeze:
        x_trg, x_src, q = x_trg[None], x_src[None], q[None]
        geom_trg = None if geom_trg is None else geom_trg[None]
        geom_src = None if geom_src is None else geom_src[None]
    B, m, d = x_trg.shape
    n = x_src.shape[1]
    from .operators import KID_HELMHOLTZ_2D, KID_HELMHOLTZ_3D, KID_HELMHOLTZ_DL_3D, KID_HELMHOLTZ_ADL_3D
    cplx_kernel = kernel_id in (KID_HELMHOLTZ_2D, KID_HELMHOLTZ_3D, KID_HELMHOLTZ_DL_3D, KID_HELMHOLTZ_ADL_3D)
    cdt = torch.complex128 if real == torch.float64 else torch.complex64
    if cplx_kernel or q.is_complex():
        qc = q.to(device=dev, dtype=cdt).contiguous()
        qr = torch.view_as_real(qc).reshape(B, n, -1)
    else:
        qr = q.to(device=dev, dtype=real).contiguous()
    width = qr.shape[-1]
    xt = x_trg.to(device=dev, dtype=real).contiguous()
    xs = x_src.to(device=dev, dtype=real).contiguous()
    gt = None if geom_trg is None else geom_trg.to(device=dev, dtype=real).contiguous()
    gs = None if geom_src is None else geom_src.to(device=dev, dtype=real).contiguous()
    out = torch.empty(B, m, width, device=dev, dtype=real)
    prec = 0 if real == torch.float64 else 1
    check(lib().skf_p2p_blocks(kernel_id, prec, width, float(param), B, m, n, d, ptr(xt), ptr(xs), ptr(gt), ptr(gs),
                               ptr(qr), ptr(out), 1 if self_interaction else 0, stream_handle(dev)))
    if cplx_kernel or q.is_complex():
        out = torch.view_as_complex(out.reshape(B, m, -1, 2))
    return out[0] if squeeze else out
