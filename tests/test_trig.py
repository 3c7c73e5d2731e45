"""The FP32 form's reduced-range sine/cosine (``sincos_reduced`` in ``csrc/p2p_core.h``).

The constants are read from the header itself and the algorithm (Cody-Waite reduction by
pi/2 in three FP32 parts, Taylor polynomials on |y| <= pi/4, quadrant selection) is run in
numpy FP32 with each FMA rounded once; the result must match sin and cos in FP64 to a few
FP32 ulps over the whole range the host allows (|x| <= 65536), so a wrong digit, a dropped
term or a swapped quadrant in the header fails here. The device instances are checked end
to end by the FP32 Helmholtz tests on the GPU.
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


def _sincos(x, k):
    x = np.float32(x)
    two_over_pi, c1, c2, c3, s9, s7, s5, s3, c10, c8, c6, c4, c2h, one = [np.float32(v) for v in k]
    n = np.float32(np.rint(np.float32(x * two_over_pi)))
    y = _fma(-n, c1, x)
    y = _fma(-n, c2, y)
    y = _fma(-n, c3, y)
    z = np.float32(y * y)
    ps = _fma(z, s9, s7)
    ps = _fma(z, ps, s5)
    ps = _fma(z, ps, s3)
    sn = _fma(np.float32(y * z), ps, y)
    pc = _fma(z, c10, c8)
    pc = _fma(z, pc, c6)
    pc = _fma(z, pc, c4)
    pc = _fma(z, pc, c2h)
    cs = _fma(z, pc, one)
    q = int(n)
    sv, cv = (cs, sn) if q & 1 else (sn, cs)
    return (-sv if q & 2 else sv), (-cv if (q + 1) & 2 else cv)


def test_constants_are_the_mathematical_ones():
    k = _consts()
    assert len(k) == 14
    getcontext().prec = 50
    pi = Decimal("3.14159265358979323846264338327950288419716939937510")
    half = pi / 2
    assert abs(Decimal(float(np.float32(k[0]))) - 2 / pi) < Decimal("1e-7")
    split = Decimal(float(np.float32(k[1]))) + Decimal(float(np.float32(k[2]))) + Decimal(float(np.float32(k[3])))
    assert abs(split - half) < Decimal("1e-21"), "pi/2 must be split exactly enough for |n| up to 4e4"
    taylor = [1 / 362880, -1 / 5040, 1 / 120, -1 / 6, -1 / 3628800, 1 / 40320, -1 / 720, 1 / 24, -0.5, 1.0]
    np.testing.assert_allclose(k[4:], taylor, rtol=1e-7)


def test_reduced_sincos_matches_fp64_over_the_allowed_range():
    k = _consts()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-4.0, 4.0, 3000), rng.uniform(-65536.0, 65536.0, 3000),
                         np.arange(-40, 41) * (np.pi / 4), [0.0, 1e-30, 65536.0, -65536.0]]).astype(np.float32)
    err = 0.0
    for x in xs:
        s, c = _sincos(x, k)
        xd = float(x)  # the FP32 input is exact in FP64
        err = max(err, abs(float(s) - np.sin(xd)), abs(float(c) - np.cos(xd)))
    # reduction rounding grows with |n| as ulp(x) * 2^-24 * |n|-ish; a few FP32 ulps of 1
    assert err < 4e-7, err


def test_quadrants_and_signs():
    k = _consts()
    for q in range(-8, 9):
        x = np.float32(q * np.pi / 2 + 0.3)
        s, c = _sincos(x, k)
        xd = float(x)
        assert abs(float(s) - np.sin(xd)) < 3e-7 and abs(float(c) - np.cos(xd)) < 3e-7
