"""Pins the oracle's sine and cosine (orc_sin / orc_cos, the rotation angle
of expRodrigues, RigidBody.java:382-401), which follow fdlibm 5.3 as
java.lang.StrictMath does: the reduction constants must add up to pi/2 to
the precision each step claims, every result must lie within fdlibm's
documented bound of one ulp of the exact value (mpmath at 200 bits) over
the kernel range, the three reduction paths and the points next to
multiples of pi/2, and the quadrant logic must give the right signs."""
import math

import numpy as np
import pytest

from oracle import pyoracle

mpmath = pytest.importorskip("mpmath")


@pytest.fixture(scope="module")
def lib():
    return pyoracle.load()


def _ulps(got, exact):
    return float(abs(mpmath.mpf(got) - exact) / math.ulp(float(exact)))


def test_within_one_ulp_of_exact(lib):
    mpmath.mp.prec = 200
    rng = np.random.default_rng(0)
    xs = np.concatenate([
        np.exp(rng.uniform(math.log(1e-9), math.log(0.78), 3000)),   # kernels only
        rng.uniform(0.78, 2.4, 2000),                                  # n = +-1 path
        rng.uniform(2.4, 2000.0, 2000),                                # medium path
        [k * math.pi / 2 for k in range(1, 64)],                       # next to n pi/2: 2nd/3rd step
        [np.nextafter(k * math.pi / 2, 0.0) for k in range(1, 64)],
        [np.nextafter(k * math.pi / 2, 10.0 * k) for k in range(1, 64)],
    ])
    worst = 0.0
    for x in xs:
        x = float(x)
        for mine, exact in ((lib.orc_sin, mpmath.sin), (lib.orc_cos, mpmath.cos)):
            for arg in (x, -x):
                worst = max(worst, _ulps(mine(arg), exact(mpmath.mpf(arg))))
    assert worst < 1.0, worst


def test_reduction_constants_split_pi_over_two():
    mpmath.mp.prec = 200
    half_pi = mpmath.pi / 2
    p1, p1t = 1.57079632673412561417e+00, 6.07710050650619224932e-11
    p2, p2t = 6.07710050630396597660e-11, 2.02226624879595063154e-21
    p3, p3t = 2.02226624871116645580e-21, 8.47842766036889956997e-32
    m = mpmath.mpf
    assert abs(m(p1) + m(p1t) - half_pi) < m(2) ** -85
    assert abs(m(p1) + m(p2) + m(p2t) - half_pi) < m(2) ** -118
    assert abs(m(p1) + m(p2) + m(p3) + m(p3t) - half_pi) < m(2) ** -150
    # the leading parts are short enough that n * part is exact up to n = 2^19 + 2^18
    n = 2 ** 19 + 2 ** 18
    for part in (p1, p2, p3):
        assert m(n) * m(part) == m(n * part)


def test_quadrants_and_special_values(lib):
    for k in range(-8, 9):
        x = k * math.pi / 2 + 0.3
        assert math.copysign(1.0, lib.orc_sin(x)) == math.copysign(1.0, math.sin(x))
        assert math.copysign(1.0, lib.orc_cos(x)) == math.copysign(1.0, math.cos(x))
    assert lib.orc_sin(0.0) == 0.0 and lib.orc_cos(0.0) == 1.0
    assert lib.orc_sin(1e-300) == 1e-300
    assert math.isnan(lib.orc_sin(math.inf)) and math.isnan(lib.orc_cos(math.nan))
    # sin(x)^2 + cos(x)^2 = 1 to a few ulps
    for x in np.linspace(0.0, 100.0, 1001):
        s, c = lib.orc_sin(float(x)), lib.orc_cos(float(x))
        assert abs(s * s + c * c - 1.0) < 4e-16
