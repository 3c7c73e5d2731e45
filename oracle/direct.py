"""ctypes wrapper of the C direct-sum restatement ``p2p_ref.c`` (oracle; test infrastructure only)."""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "p2p_ref.c")
LIB = os.path.join(HERE, "liboracle_p2p.so")

KERNEL_IDS = {
    "laplace_2d": 0,
    "laplace_3d": 1,
    "helmholtz_2d": 2,
    "helmholtz_3d": 3,
    "laplace_double_layer_3d": 4,
    "laplace_adjoint_double_layer_3d": 5,
    "helmholtz_double_layer_3d": 6,
    "helmholtz_adjoint_double_layer_3d": 7,
}

_lib = None


def build():
    """Compile the C oracle with gcc (OpenMP)."""
    subprocess.run(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", SRC, "-o", LIB, "-lm"], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        f = L.oracle_direct_sum
        f.restype = ctypes.c_int
        vp = ctypes.c_void_p
        f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, vp, ctypes.c_long, vp, vp, vp, vp, ctypes.c_int,
                      ctypes.c_double, vp]
        _lib = L
    return _lib


def direct_sum_c(name, X, Y, q, kappa=0.0, geom_trg=None, geom_src=None):
    """``sum_j K(X_i, Y_j) q_j`` in FP64 (C, OpenMP); ``q`` real or complex, 1D or 2D."""
    kid = KERNEL_IDS[name]
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    q = np.asarray(q)
    squeeze = q.ndim == 1
    if squeeze:
        q = q[:, None]
    cplx = kid in (2, 3, 6, 7) or np.iscomplexobj(q)
    if cplx:
        qc = np.ascontiguousarray(q, dtype=np.complex128)
        qr = qc.view(np.float64).reshape(q.shape[0], -1)
    else:
        qr = np.ascontiguousarray(q, dtype=np.float64)
    width = qr.shape[1]
    out = np.zeros((X.shape[0], width))
    gt = None if geom_trg is None else np.ascontiguousarray(geom_trg, dtype=np.float64)
    gs = None if geom_src is None else np.ascontiguousarray(geom_src, dtype=np.float64)
    rc =_load().oracle_direct_sum(kid, X.shape[1], X.shape[0], X.ctypes.data, Y.shape[0], Y.ctypes.data,
                                   None if gt is None else gt.ctypes.data, None if gs is None else gs.ctypes.data,
                                   qr.ctypes.data, width, float(kappa), out.ctypes.data)
    if rc != 0:
        raise ValueError("oracle_direct_sum rejected its arguments")
    if cplx:
        out = out.view(np.complex128).reshape(X.shape[0], -1)
    return out[:, 0] if squeeze else out
