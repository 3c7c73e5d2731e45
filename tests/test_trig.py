"""The FP32 form's reduced-range sine/cosine (``sincos_reduced`` in ``csrc/p2p_core.h``).

The constants are read from the header itself and the reduction (Cody-Waite by 2 pi in three
FP32 parts, then y / (2 pi) revolutions for the hardware sine and cosine) is run in numpy FP32
with each FMA rounded once: the revolutions handed to the hardware must equal the exact
x / (2 pi) minus an integer to a few FP32 ulps of 1/2 over the whole range the host allows
(|x| <= 65536), so a wrong digit, a dropped part of the split or a wrong scale fails here. The
hardware sine/cosine itself (v_sin_f32 / v_cos_f32, measured max absolute error 1.3e-7 over
[-1/2, 1/2] revolutions against FP64 on an MI355X) is checked end to end on the GPU
(``test_gpu_parity.py::test_fp32_helmholtz_fast_trig_accuracy``).
"""

import os
import re
from decimal import Decimal, getcontext

import numpy as np

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "skelfmm_amd", "csrc",
                      "p2p_core.h")


def _body():
    src = open(HEADER, encoding="utf-8").read()
    a = src.index("__device__ inline void sincos_reduced(")
    b = src.index("\n}\n", a)
    return src[a:b]


def _consts():
    return [float(m) for m in re.findall(r"(-?\d+\.\d+(?:e[-+]?\d+)?)f", _body())]


def _fma(a, b, c):
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def _revolutions(x, k):
    inv1, c1, c2, c3, inv2 = [np.float32(v) for v in k]
    x = np.float32(x)
    n = np.float32(np.rint(np.float32(x * inv1)))
    y = _fma(-n, c1, x)
    y = _fma(-n, c2, y)
    y = _fma(-n, c3, y)
    return np.float32(y * inv2), n


def test_constants_are_the_mathematical_ones():
    k = _consts()
    assert len(k) == 5
    getcontext().prec = 50
    pi = Decimal("3.14159265358979323846264338327950288419716939937510")
    two_pi = 2 * pi
    for inv in (k[0], k[4]):
        assert abs(Decimal(float(np.float32(inv))) - 1 / two_pi) < Decimal("1e-8")
    split = sum(Decimal(float(np.float32(c))) for c in k[1:4])
    assert abs(split - two_pi) < Decimal("1e-21"), "2 pi must be split exactly enough for |n| up to 1.1e4"
    assert float(np.float32(k[1])) == float(np.float32(2 * np.pi))


def test_reduction_matches_exact_revolutions_over_the_allowed_range():
    k = _consts()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-8.0, 8.0, 3000), rng.uniform(-65536.0, 65536.0, 3000),
                         np.arange(-40, 41) * (np.pi / 4), [0.0, 1e-30, 65536.0, -65536.0]]).astype(np.float32)
    getcontext().prec = 40
    two_pi = 2 * Decimal("3.14159265358979323846264338327950288419716939937510")
    worst = 0.0
    for x in xs:
        t, n = _revolutions(x, k)
        exact = Decimal(float(x)) / two_pi - Decimal(float(n))  # the FP32 input is exact
        worst = max(worst, abs(float(Decimal(float(t)) - exact)))
        # n = rint(fl(x / 2 pi)) may pick the neighbour near a half turn: |t| <= 1/2 + |x| 2^-24 / (2 pi)
        assert abs(float(t)) <= 0.5 + 1e-3
        # with an ideal sine of the revolutions the value is within the reduction error
        assert abs(np.sin(2 * np.pi * float(t)) - np.sin(float(x))) < 4e-7
    assert worst < 6e-8, worst


def test_quadrants_and_signs():
    k = _consts()
    for q in range(-8, 9):
        x = np.float32(q * np.pi / 2 + 0.3)
        t, _ = _revolutions(x, k)
        xd = float(x)
        assert abs(np.sin(2 * np.pi * float(t)) - np.sin(xd)) < 3e-7
        assert abs(np.cos(2 * np.pi * float(t)) - np.cos(xd)) < 3e-7
