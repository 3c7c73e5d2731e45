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


# ------------------------------------------------------------------ FP64 (the plain form)
def _body64():
    src = open(HEADER, encoding="utf-8").read()
    a = src.index("__device__ inline void sincos_reduced(double x")
    b = src.index("\n}\n", a)
    return src[a:b]


def _consts64():
    return [float(m) for m in re.findall(r"(?<![\w.])(-?\d+\.\d+(?:e[-+]?\d+)?)(?![\w.])", _body64())]


def _fma64(a, b, c):
    from fractions import Fraction

    return float(Fraction(a) * Fraction(b) + Fraction(c))  # exact, then one rounding


def _sincos64(x, k):
    two_over_pi, c1, c2 = k[:3]
    sin_c = k[3:10]   # y^15 .. y^3 coefficients
    cos_c = k[10:17]  # y^16 .. y^4
    c2h, one = k[17], k[18]
    n = float(round(x * two_over_pi))  # rint: round half to even
    y = _fma64(-n, c1, x)
    y = _fma64(-n, c2, y)
    z = y * y
    ps = _fma64(z, sin_c[0], sin_c[1])
    for c in sin_c[2:]:
        ps = _fma64(z, ps, c)
    sn = _fma64(y * z, ps, y)
    pc = _fma64(z, cos_c[0], cos_c[1])
    for c in cos_c[2:] + [c2h]:
        pc = _fma64(z, pc, c)
    cs = _fma64(z, pc, one)
    q = int(n)
    sv, cv = (cs, sn) if q & 1 else (sn, cs)
    return (-sv if q & 2 else sv), (-cv if (q + 1) & 2 else cv)


def test_fp64_constants_are_the_mathematical_ones():
    import math

    k = _consts64()
    assert len(k) == 19
    getcontext().prec = 60
    pi = Decimal("3.14159265358979323846264338327950288419716939937510582097494459")
    assert k[0] == float(2 / pi)
    assert k[1] == float(pi / 2) and k[2] == float(pi / 2 - Decimal(k[1]))
    sin_ref = [(-1) ** j / Decimal(math.factorial(2 * j + 1)) for j in range(7, 0, -1)]
    cos_ref = [(-1) ** j / Decimal(math.factorial(2 * j)) for j in range(8, 0, -1)]
    assert k[3:10] == [float(v) for v in sin_ref]
    assert k[10:18] == [float(v) for v in cos_ref] and k[18] == 1.0


def test_fp64_sincos_matches_libm_over_the_allowed_range():
    import math

    k = _consts64()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-8.0, 8.0, 400), rng.uniform(-65536.0, 65536.0, 400),
                         np.arange(-20, 21) * (np.pi / 4), [0.0, 1e-300, 65536.0, -65536.0, 1e-8]])
    err = 0.0
    for x in xs:
        s, c = _sincos64(float(x), k)
        err = max(err, abs(s - math.sin(float(x))), abs(c - math.cos(float(x))))
    # a few ulps of 1 (reduction ~1.1e-16, polynomial ~1 ulp, libm reference <= 1 ulp)
    assert err < 5e-16, err
