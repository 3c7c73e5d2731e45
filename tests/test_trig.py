"""The device sine/cosine of the Helmholtz kernels (``csrc/p2p_core.h``).

FP32 (``sincos_rev``): instances whose arguments the host bounds (kappa r <= kTrigDirectMax) take
kappa / (2 pi) as their constant, form the phase in revolutions t = r kappa / (2 pi) and hand it
to the hardware sine and cosine, which take revolutions and reduce them exactly within their
input range |t| <= 256 (v_sin_f32 / v_cos_f32; measured max absolute error 1.25e-7 against FP64
on an MI355X for every |t| up to 300, checked end to end on the GPU by
``test_gpu_parity.py::test_fp32_helmholtz_fast_trig_accuracy``). Here, in numpy FP32: the
constants are 1/(2 pi) and 2 pi; the host's bound keeps |t| inside the hardware range; and the
phase error of the whole path (constant and product each rounded once) stays within two
roundings of the exact kappa r, as forming kappa r in FP32 does.

FP64 (``sincos_reduced(double)``): Cody-Waite reduction by pi/2 and Taylor polynomials, run with
exact FMAs against libm.
"""

import os
import re
from decimal import Decimal, getcontext

import numpy as np

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "skelfmm_amd", "csrc",
                      "p2p_core.h")
TRIG_FAST_MAX = 65536.0  # FP64 instances
TRIG_DIRECT_MAX = 1600.0  # FP32 instances


def _const(name):
    src = open(HEADER, encoding="utf-8").read()
    m = re.search(r"constexpr (?:float|double) " + name + r" = (-?\d+\.\d+(?:e[-+]?\d+)?)f?;", src)
    assert m, name
    return float(m.group(1))


def _phase_rev(r, kappa):
    """The device path in FP32: constant kappa / (2 pi) rounded once, product rounded once;
    returns the phase in revolutions (the hardware sine/cosine reduce it themselves)."""
    krev = np.float32(np.float64(kappa) * _const("kInvTwoPi"))
    return np.float32(np.float32(r) * krev)


def test_constants_are_the_mathematical_ones():
    getcontext().prec = 50
    pi = Decimal("3.14159265358979323846264338327950288419716939937510")
    assert abs(Decimal(_const("kInvTwoPi")) - 1 / (2 * pi)) < Decimal("1e-17")
    assert _const("kTwoPiF") == float(2 * pi) and np.float32(_const("kTwoPiF")) == np.float32(2 * np.pi)
    assert _const("kTrigFastMax") == TRIG_FAST_MAX
    assert _const("kTrigDirectMax") == TRIG_DIRECT_MAX


def test_phase_stays_inside_the_hardware_range():
    """At the host's bound every phase the FP32 path forms, roundings included, is below 256
    revolutions (the input range of v_sin_f32 / v_cos_f32), for any kappa."""
    rng = np.random.default_rng(0)
    for kappa in (1e-3, 1.0, 10 * np.pi, 40.0, 1600.0, 1e5):
        rmax = np.float32(TRIG_DIRECT_MAX / kappa)
        r = np.concatenate([rng.uniform(0.0, rmax, 2000), [rmax, np.nextafter(rmax, np.float32(0))]]).astype(np.float32)
        t = _phase_rev(r, kappa)
        assert np.all(np.abs(t) < 255.0), float(np.max(np.abs(t)))


def test_phase_error_is_two_roundings_of_kappa_r():
    """Against the exact phase kappa r (FP64 for the FP32 inputs r, kappa): at most two FP32
    roundings of kappa r plus the constant's, over the range the host allows, and never worse
    than forming kappa r in FP32 and reducing it exactly."""
    rng = np.random.default_rng(1)
    worst_new = worst_ref = 0.0
    for kappa in (1.0, 10 * np.pi, 20.0, 40.0, 1000.0):
        rmax = TRIG_DIRECT_MAX / kappa
        r = np.concatenate([rng.uniform(0.0, min(rmax, 1.0), 2000), rng.uniform(0.0, rmax, 2000)]).astype(np.float32)
        y = _phase_rev(r, kappa)
        exact = (r.astype(np.float64) * np.float64(np.float32(kappa))) / (2 * np.pi)  # revolutions
        err = np.abs(((y.astype(np.float64) - exact) + 0.5) % 1.0 - 0.5) * 2 * np.pi  # radians
        bound = 3 * 2.0**-24 * exact * 2 * np.pi + 1e-12
        assert np.all(err <= bound), float(np.max(err / bound))
        kr = (r * np.float32(kappa)).astype(np.float32)  # kappa r formed in FP32, reduced exactly
        err_ref = np.abs(((kr.astype(np.float64) / (2 * np.pi) - exact) + 0.5) % 1.0 - 0.5) * 2 * np.pi
        worst_new = max(worst_new, float(np.max(err / np.maximum(exact, 1.0))))
        worst_ref = max(worst_ref, float(np.max(err_ref / np.maximum(exact, 1.0))))
        # with an ideal sine of the revolutions, sin and cos follow the phase
        assert np.max(np.abs(np.sin(2 * np.pi * y.astype(np.float64)) - np.sin(2 * np.pi * exact))) <= np.max(bound) + 1e-9
    assert worst_new <= 2.0 * worst_ref + 1e-12, (worst_new, worst_ref)


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
