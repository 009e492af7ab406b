"""The two division- and pow-free forms the device TENO5 uses, checked
against what they stand for (CPU, exact rational arithmetic):

* the cutoff test  g - P < E - 2^-70 gs  (P + E = 1e-5 gs exactly) decides
  RN(g / gs) < 1e-5 exactly as the reference's division does;
* the compensated sixth power returns the correctly rounded x^6.

FMA is emulated exactly with fractions (float(Fraction) rounds correctly),
so these are the same operations the device performs."""
import math
import random
from fractions import Fraction

import numpy as np

CT = 1.0e-5
HALF_GAP = 2.0 ** -70


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def teno_cut(g, gs):
    P = CT * gs
    E = fma(CT, gs, -P)
    return g - P < E - HALF_GAP * gs


def pow6(x):
    p2 = x * x
    e2 = fma(x, x, -p2)
    p3 = p2 * x
    e3 = fma(p2, x, -p3) + e2 * x
    p6 = p3 * p3
    e6 = fma(p3, p3, -p6) + 2.0 * p3 * e3
    return p6 + e6


def test_half_gap_is_half_the_ulp_below_ct():
    assert (CT - np.nextafter(CT, 0.0)) / 2 == HALF_GAP


def test_cutoff_matches_division_near_and_far_from_the_threshold():
    rng = random.Random(11)
    cases = 0
    for _ in range(4000):
        gs = math.exp(rng.uniform(0.0, 40.0))
        # quotients within a few ulps of CT, and far from it
        for k in range(-6, 7):
            q = CT * (1.0 + k * 2.0 ** -52) * (1.0 + rng.uniform(-1, 1) * 2.0 ** -50)
            g = q * gs
            assert teno_cut(g, gs) == (g / gs < CT), (g, gs)
            cases += 1
        g = gs * math.exp(rng.uniform(-60.0, 0.0))
        assert teno_cut(g, gs) == (g / gs < CT)
    assert cases > 50000


def test_pow6_is_correctly_rounded():
    rng = random.Random(5)
    agree_libm = 0
    n = 3000
    for _ in range(n):
        x = math.exp(rng.uniform(0.0, 50.0))
        exact = float(Fraction(x) ** 6)
        if math.isinf(exact):
            continue
        assert pow6(x) == exact, x
        agree_libm += math.pow(x, 6.0) == exact
    # libm's pow is within its last ulp of the same value
    assert agree_libm > 0.9 * n
