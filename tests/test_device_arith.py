"""The division- and pow-free forms the device TENO5 uses, checked against
what they stand for (CPU, exact rational arithmetic where it matters):

* the cutoff test  g - P < E - 2^-70 gs  (P + E = 1e-5 gs exactly) decides
  RN(g / gs) < 1e-5 exactly as the reference's division does;
* the stencil selection from gamma_k = (1 + tau5 / e_k)^6 with the three
  quotients taken from one reciprocal of e0 e1 e2 and the sixth power as
  three products (sf_face.h teno5_weights) keeps the reference's stencils
  (SF_TENO5.cpp:67-75: a correctly rounded quotient, std::pow) on data
  over twelve decades, and differs from them only where gamma_k / sum
  lies within rounding of the cutoff.

FMA is emulated exactly with fractions (float(Fraction) rounds correctly),
so the cutoff test performs the same operations as the device."""
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


def device_selection(b):
    """teno5_weights of sf_face.h: the quotients from one reciprocal (the
    device refines the hardware estimate to within about an ulp; here the
    correctly rounded reciprocal), the sixth power as three products."""
    tau5 = abs(b[0] - b[2])
    e = [x + 1.0e-40 for x in b]
    e01 = e[0] * e[1]
    t = tau5 * (1.0 / (e01 * e[2]))
    g = []
    # 1 + t e contracts to one fused multiply-add
    for x in (fma(t, e[1] * e[2], 1.0), fma(t, e[0] * e[2], 1.0), fma(t, e01, 1.0)):
        x3 = x * x * x
        g.append(x3 * x3)
    gs = g[0] + g[1] + g[2]
    return [not teno_cut(gk, gs) for gk in g], g, gs


def reference_selection(b):
    """SF_TENO5.cpp:67-75 as the reference evaluates it."""
    def pow6(x):
        try:
            return math.pow(x, 6.0)
        except OverflowError:      # C's pow returns inf
            return math.inf

    tau5 = abs(b[0] - b[2])
    g = [pow6(1.0 + tau5 / (x + 1.0e-40)) for x in b]
    gs = g[0] + g[1] + g[2]
    return [not (gk / gs < CT) for gk in g], g, gs


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


def test_selection_keeps_the_reference_stencils():
    rng = random.Random(5)
    cut = 0
    for _ in range(60000):
        # smoothness indicators over twelve decades, one stencil often far
        # rougher than the others (where cutoffs happen)
        b = [math.exp(rng.uniform(-28.0, 0.0)) for _ in range(3)]
        b[rng.randrange(3)] *= math.exp(rng.uniform(0.0, 30.0))
        if rng.random() < 0.1:
            b[rng.randrange(3)] = 0.0
        ref, g, gs = reference_selection(b)
        if not math.isfinite(gs):
            continue   # a fault on both sides (requireFiniteGammaSum)
        mine, _, _ = device_selection(b)
        cut += not all(ref)
        assert mine == ref, b
    assert cut > 5000   # the cutoff is exercised, not just the all-kept case


def test_selection_differs_only_within_rounding_of_the_cutoff():
    """Indicators placed where gamma_2 / sum crosses 1e-5 (found by bisection
    on beta_2 with the reference's formula), then moved by a few ulps: where
    the device's selection differs from the reference's, the reference's
    ratio is within 1e-14 (relative) of the cutoff."""
    rng = random.Random(9)
    near = differ = 0
    for _ in range(300):
        b0 = math.exp(rng.uniform(-20.0, 0.0))
        b1 = b0 * math.exp(rng.uniform(-1.0, 1.0))

        def ratio(b2):
            _, g, gs = reference_selection([b0, b1, b2])
            return g[2] / gs

        lo, hi = b0, b0 * 1e6            # ratio(lo) > CT > ratio(hi)
        if not (ratio(lo) > CT > ratio(hi)):
            continue
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if mid in (lo, hi):
                break
            lo, hi = (mid, hi) if ratio(mid) > CT else (lo, mid)
        for k in range(-40, 41):
            b2 = hi * (1.0 + k * 2.0 ** -52)
            b = [b0, b1, b2]
            ref, g, gs = reference_selection(b)
            mine, _, _ = device_selection(b)
            near += abs(g[2] / gs - CT) <= 1e-14 * CT
            if mine != ref:
                differ += 1
                assert abs(g[2] / gs - CT) <= 1e-14 * CT, (b, g[2] / gs)
    assert near > 2000     # the samples do sit at the cutoff
