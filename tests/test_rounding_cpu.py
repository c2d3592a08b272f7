"""Functions whose last bit the reference leaves to its math library, and
which the oracle therefore fixes at the correctly rounded value:

* x**1.5 of the viscosity law mu = (1 + S) T**1.5 / (T + S) / Re
  (src/Main3D.f90:582).  Correct rounding is decided in exact rational
  arithmetic: y is the nearest double to x^1.5 exactly when the midpoints to
  its neighbours bracket it, m_lo^2 <= x^3 <= m_hi^2.  The unrefined
  x*sqrt(x) fails that check on a share of the same samples (so the check
  discriminates), and the C library's pow stays within one unit in the last
  place of the oracle's value.
* sin and cos of the per-stage scalars: the moving frame
  (src/Subroutineso.f90:192-259), the time-step ramp (src/Main3D.f90:610) and
  the body force (src/rpt.f90:761-763).  Checked against 300-bit mpmath values:
  the distance to the exact value is at most half the gap to the neighbour on
  that side; libm's sin and cos are one unit off on a few of the samples.
"""
import math
from fractions import Fraction

import mpmath
import numpy as np

import pyoracle as po


def correctly_rounded(x, y):
    if y == 0 or not math.isfinite(y):
        return False
    lo = (Fraction(y) + Fraction(math.nextafter(y, 0.0))) / 2
    hi = (Fraction(y) + Fraction(math.nextafter(y, math.inf))) / 2
    x3 = Fraction(x) ** 3
    return lo * lo <= x3 <= hi * hi


def samples():
    rng = np.random.default_rng(15)
    xs = list(rng.uniform(0.5, 2.0, 4000))                 # gamma p / rho near the free stream
    xs += list(np.exp(rng.uniform(np.log(1e-6), np.log(1e6), 3000)))
    xs += [float(k * k) / 2 ** 20 for k in range(1, 600, 7)]  # perfect squares
    xs += [2.0 ** e for e in range(-40, 41)]
    one = 1.0
    for _ in range(50):                                    # neighbours of 1
        xs.append(one)
        one = math.nextafter(one, 2.0)
    return [float(x) for x in xs]


def test_pow15_is_correctly_rounded():
    naive_wrong = 0
    libm_diff = 0
    for x in samples():
        y = po.pow15(x)
        assert correctly_rounded(x, y), (x, y)
        if not correctly_rounded(x, x * math.sqrt(x)):
            naive_wrong += 1
        p = math.pow(x, 1.5)
        if p != y:
            libm_diff += 1
            assert p in (math.nextafter(y, 0.0), math.nextafter(y, math.inf)), (x, p, y)
    # one unrefined rounding too many: x*sqrt(x) is off on a sizeable share
    assert naive_wrong > 100, naive_wrong
    print(f"x*sqrt(x) off by one unit on {naive_wrong} samples; libm pow on {libm_diff}")


def test_pow15_special_values():
    assert po.pow15(4.0) == 8.0 and po.pow15(0.25) == 0.125
    assert po.pow15(0.0) == 0.0
    assert math.isnan(po.pow15(-1.0)) and math.isnan(po.pow15(float("nan")))
    assert po.pow15(math.inf) == math.inf
    for x in (1e-200, 3e-95, 2e95, 7e150):                  # outside the refined range: pow's value
        assert po.pow15(x) == math.pow(x, 1.5)


def nearest(y, exact):
    """y is the double nearest the mpmath value `exact`."""
    d = mpmath.mpf(y) - exact
    nb = math.nextafter(y, math.inf) if d < 0 else math.nextafter(y, -math.inf)
    return abs(d) <= abs(mpmath.mpf(nb) - mpmath.mpf(y)) / 2


def test_sincos_is_correctly_rounded():
    mpmath.mp.prec = 300   # before the samples: they include mpmath's k pi/2
    rng = np.random.default_rng(27)
    xs = list(rng.uniform(0.0, 4.0, 3000))                 # frame and ramp arguments (0 .. pi)
    xs += list(rng.uniform(-2e3, 2e3, 2000))               # body-force phases, many quadrants
    xs += [0.05 * math.pi * k for k in range(10)]           # the ramp's ten arguments
    xs += [k * math.pi / 2 + d for k in range(-8, 9) for d in (0.0, -1e-9, 1e-12, 3e-6)]   # at and near quadrant ends
    xs += [1e-300, 2.0 ** -30, 0.7853981633974483, 1e5]
    # the doubles nearest k pi/2: the result of sin or cos is then ~1e-16 x,
    # and the last part of pi/2 in the reduction decides its rounding
    xs += [float(mpmath.mpf(k) * mpmath.pi / 2) for k in range(1, 400)]
    libm = 0
    for x in xs:
        x = float(x)
        s, c = po.sincos(x)
        X = mpmath.mpf(x)
        assert nearest(s, mpmath.sin(X)), ("sin", x, s)
        assert nearest(c, mpmath.cos(X)), ("cos", x, c)
        libm += (math.sin(x) != s) + (math.cos(x) != c)
    print(f"libm sin/cos differ from the correctly rounded value on {libm} of {2 * len(xs)}")


def test_sincos_special_values():
    assert po.sincos(0.0) == (0.0, 1.0)
    assert po.sincos(2.0 ** -30) == (2.0 ** -30, 1.0)
    for x in (2.0 ** 21, -3e7, float("inf"), float("nan")):     # outside the reduced range: libm's values
        s, c = po.sincos(x)
        assert (s == math.sin(x) or (math.isnan(s) and math.isnan(math.sin(x)))) if math.isfinite(x) else math.isnan(s)
