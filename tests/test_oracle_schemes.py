"""Pins of the oracle's WENO3 / WENO7 reconstructions and its
Steger-Warming, Lax-Friedrichs and Roe fluxes (oracle/sfcpu.c), checked
against what mathematics fixes: polynomial reproduction in exact rational
arithmetic, the closed-form smoothness integral, consistency and upwinding
of the flux splittings, and a textbook wave decomposition of the Roe flux.
The end-to-end pin is the reference's own frozen WENO7 + Rusanov + forward
Euler Sod run (test/Sod/sodCase_weno7_t0p2/baseline.json, copied into
tests/golden/sod_weno7_t0p2/)."""
import json
import math
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from oracle import bind as ob

GAMMA = 1.4
GOLDEN = Path(__file__).resolve().parent / "golden" / "sod_weno7_t0p2" / "baseline.json"


def prim(rho, u, v, w, p, g=GAMMA):
    return np.array([rho, rho * u, rho * v, rho * w,
                     p / (g - 1.0) + 0.5 * rho * (u * u + v * v + w * w)])


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


# ---- exact-arithmetic helpers ------------------------------------------------

def cell_average_basis(m):
    """h_m with cell averages x^m: the polynomial whose average over
    [s-1/2, s+1/2] equals s^m for every integer s (exact, by solving the
    triangular system of averaged monomials)."""
    # average of x^j over [s-1/2, s+1/2] = ((s+1/2)^(j+1) - (s-1/2)^(j+1)) / (j+1)
    # as a polynomial in s, whose coefficients fill column j
    size = m + 1
    A = [[Fraction(0)] * size for _ in range(size)]
    for j in range(size):
        for k in range(j + 2):
            c = Fraction(math.comb(j + 1, k)) * (Fraction(1, 2) ** (j + 1 - k))
            c = c - Fraction(math.comb(j + 1, k)) * (Fraction(-1, 2) ** (j + 1 - k))
            if k < size:
                A[k][j] += c / (j + 1)
    # solve A h = e_m (upper triangular: coefficient of s^k needs j >= k)
    h = [Fraction(0)] * size
    for k in range(size - 1, -1, -1):
        rhs = Fraction(1 if k == m else 0) - sum(A[k][j] * h[j] for j in range(k + 1, size))
        h[k] = rhs / A[k][k]
    return h


def poly_at(coef, x):
    return sum(c * Fraction(x) ** k for k, c in enumerate(coef))


def test_cell_average_basis_is_right():
    # x^2 -> x^2 - 1/12 and x^3 -> x^3 - x/4 (the forms used for WENO5)
    assert cell_average_basis(2) == [Fraction(-1, 12), 0, 1]
    assert cell_average_basis(3) == [0, Fraction(-1, 4), 0, 1]


# WENO7 candidates and linear weights as the reference defines them
# (SF_WENO7.cpp:73-78), in exact arithmetic
W7_CAND = [
    [Fraction(-3, 12), Fraction(13, 12), Fraction(-23, 12), Fraction(25, 12), 0, 0, 0],
    [0, Fraction(1, 12), Fraction(-5, 12), Fraction(13, 12), Fraction(3, 12), 0, 0],
    [0, 0, Fraction(-1, 12), Fraction(7, 12), Fraction(7, 12), Fraction(-1, 12), 0],
    [0, 0, 0, Fraction(3, 12), Fraction(13, 12), Fraction(-5, 12), Fraction(1, 12)],
]
W7_D = [Fraction(1, 35), Fraction(12, 35), Fraction(18, 35), Fraction(4, 35)]


def test_weno7_candidates_are_fourth_order_and_weights_seventh_order():
    # node offsets -3..3; the face is x = 1/2
    for m in range(7):
        h = cell_average_basis(m)
        v = [Fraction(s) ** m for s in range(-3, 4)]
        face = poly_at(h, Fraction(1, 2))
        cand = [sum(c * x for c, x in zip(row, v)) for row in W7_CAND]
        if m <= 3:
            assert all(c == face for c in cand), m
        assert sum(d * c for d, c in zip(W7_D, cand)) == face, m
    # and the candidates of the oracle are these: on data where every
    # smoothness indicator is equal (linear data) the weights are linear
    v = [2.0 + 0.5 * s for s in range(-3, 4)]
    r, e = ob.weno7(v)
    assert e == 0 and abs(r - (2.0 + 0.25)) < 1e-15


def test_weno7_reproduces_cubics_and_is_eno_at_a_step():
    for m in range(4):
        h = cell_average_basis(m)
        v = [float(Fraction(s) ** m) for s in range(-3, 4)]
        r, e = ob.weno7(v)
        assert e == 0 and abs(r - float(poly_at(h, Fraction(1, 2)))) < 1e-12, m
    # constant data to the last bit (the candidates divide by 12)
    assert abs(ob.weno7([3.25] * 7)[0] - 3.25) <= 2 ** -52 * 3.25
    # a step between nodes 0 and 1: the all-left candidate dominates
    r, _ = ob.weno7([0, 0, 0, 0, 1, 1, 1])
    assert 0.0 <= r < 1e-9
    r, _ = ob.weno7([1, 1, 1, 1, 0, 0, 0])
    assert abs(r - 1.0) < 1e-9
    _, e = ob.weno7([0, 0, 0, float("inf"), 0, 0, 0])
    assert e & 32


def test_weno7_linear_weight_limit():
    # beta << eps: the seventh-order linear combination, exact for the cell
    # averages of a sextic
    # averages of a sextic; a wrong candidate or linear weight moves the
    # result by O(amp) = 1e-10, rounding by ~1e-16
    amp = 1e-10
    h = cell_average_basis(6)
    v = [1.0 + amp * float(Fraction(s) ** 6) for s in range(-3, 4)]
    r, _ = ob.weno7(v)
    assert abs(r - (1.0 + amp * float(poly_at(h, Fraction(1, 2))))) < 2e-15


def exact_smoothness(v, a, b):
    """int_a^b p'^2 + p''^2 + p'''^2 for the cubic through (x, v_x), x = 0..3
    (Lagrange form, exact)."""
    xs = [Fraction(x) for x in range(4)]
    vs = [Fraction(x) for x in v]
    # monomial coefficients of the interpolant
    coef = [Fraction(0)] * 4
    for i in range(4):
        basis = [Fraction(1)]
        den = Fraction(1)
        for j in range(4):
            if j == i:
                continue
            basis = [Fraction(0)] + basis
            for k in range(len(basis) - 1):
                basis[k] -= xs[j] * basis[k + 1]
            den *= xs[i] - xs[j]
        for k in range(4):
            coef[k] += vs[i] * basis[k] / den

    def deriv(c):
        return [k * c[k] for k in range(1, len(c))]

    def mul(p, q):
        out = [Fraction(0)] * (len(p) + len(q) - 1)
        for i, x in enumerate(p):
            for j, y in enumerate(q):
                out[i + j] += x * y
        return out

    def integ(c):
        A, B = Fraction(a), Fraction(b)
        return sum(ck * (B ** (k + 1) - A ** (k + 1)) / (k + 1) for k, ck in enumerate(c))

    d1 = deriv(coef)
    d2 = deriv(d1)
    d3 = deriv(d2)
    return integ(mul(d1, d1)) + integ(mul(d2, d2)) + integ(mul(d3, d3))


def test_weno7_smoothness_is_the_closed_form_integral():
    rng = np.random.default_rng(7)
    for a, b in ((2.5, 3.5), (1.5, 2.5), (0.5, 1.5), (-0.5, 0.5)):
        for _ in range(50):
            v = rng.integers(-20, 21, size=4).astype(float)
            ex = float(exact_smoothness(v, a, b))
            got = ob.weno7_smoothness(v, a, b)
            assert abs(got - ex) <= 1e-12 * max(1.0, abs(ex)), (v, a, b, got, ex)


def test_weno3_candidates_and_weights():
    # SF_WENO3.cpp:45-51: the two candidates reproduce quadratic cell averages
    for m in range(3):
        h = cell_average_basis(m)
        face = poly_at(h, Fraction(1, 2))
        v = [Fraction(s) ** m for s in range(-1, 3)]
        p0 = (-v[0] + 5 * v[1] + 2 * v[2]) / 6
        p1 = (2 * v[1] + 5 * v[2] - v[3]) / 6
        assert p0 == face and p1 == face
        r, e = ob.weno3([float(x) for x in v])
        assert e == 0 and abs(r - float(face)) < 1e-13
    # beta_k = squared spread of the candidate: with beta0 = 0 and beta1 > 0
    # the weights are (1/3)/1e-12 : (2/3)/(1e-6 + beta1)^2
    v = [1.0, 1.0, 1.0, 4.0]
    r, _ = ob.weno3(v)
    a0 = (1 / 3) / (1e-6) ** 2
    a1 = (2 / 3) / (1e-6 + 9.0) ** 2
    p0 = 1.0
    p1 = (2 + 5 - 4.0) / 6.0
    assert abs(r - (a0 * p0 + a1 * p1) / (a0 + a1)) < 1e-15
    assert ob.weno3([0.75] * 4)[0] == 0.75


# ---- face fluxes -------------------------------------------------------------

def _states(rng, count, speed=3.0):
    out = []
    for _ in range(count):
        rho = math.exp(rng.uniform(-1.5, 1.0))
        p = math.exp(rng.uniform(-1.5, 1.5))
        u, v, w = rng.uniform(-speed, speed, size=3)
        out.append(prim(rho, u, v, w, p))
    return out


NORMALS = [np.array(n, float) for n in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (0.6, 0.8, 0.0),
                                         (0.48, 0.6, 0.64))]


@pytest.mark.parametrize("flux", ["steger", "laxfriedrichs", "roe"])
def test_splitting_fluxes_are_consistent(flux):
    rng = np.random.default_rng(11)
    for q in _states(rng, 60):
        for n in NORMALS:
            rc, f = ob.face_flux(flux, q, q, n)
            rc2, fp = ob.physical_flux(q, n)
            assert rc == 0 and rc2 == 0
            assert rel(f, fp) < 1e-12, (flux, q, n)


def test_consistency_holds_inside_the_steger_warming_smoothing_band():
    # |lambda| < 2 eps_sw = 0.02 for the convective eigenvalue: the smoothed
    # split still sums to the eigenvalue
    for u in (0.0, 0.004, -0.013, 0.0199):
        q = prim(1.1, u, 0.2, -0.1, 0.9)
        rc, f = ob.face_flux("steger", q, q, np.array([1.0, 0.0, 0.0]))
        _, fp = ob.physical_flux(q, np.array([1.0, 0.0, 0.0]))
        assert rc == 0 and rel(f, fp) < 1e-13


def test_steger_warming_upwinds_supersonic_flow():
    # every eigenvalue above 0.02: F+ = F, F- = 0, so the face flux is the
    # physical flux of the left state whatever the right state is
    n = np.array([1.0, 0.0, 0.0])
    qL = prim(1.0, 3.0, 0.3, -0.2, 1.0)           # c = 1.18, u - c > 0
    qR = prim(0.4, 2.5, -0.1, 0.5, 0.6)
    rc, f = ob.face_flux("steger", qL, qR, n)
    _, fL = ob.physical_flux(qL, n)
    assert rc == 0 and rel(f, fL) < 1e-13
    rc, f = ob.face_flux("steger", -1 * qL * np.array([-1, 1, 1, 1, -1]), qR, -n)
    assert rc == 0


def test_lax_friedrichs_is_the_local_split():
    # F = (F_L + F_R)/2 + (lambda_L q_L - lambda_R q_R)/2 with each side's
    # own |u_n| + c (not the larger of the two, which would be Rusanov)
    n = np.array([0.0, 1.0, 0.0])
    qL = prim(1.0, 0.2, 0.5, 0.0, 1.0)
    qR = prim(0.5, -0.1, -0.4, 0.3, 0.4)
    _, fL = ob.physical_flux(qL, n)
    _, fR = ob.physical_flux(qR, n)
    lamL = abs(0.5) + math.sqrt(GAMMA * 1.0 / 1.0)
    lamR = abs(-0.4) + math.sqrt(GAMMA * 0.4 / 0.5)
    want = 0.5 * (fL + lamL * qL) + 0.5 * (fR - lamR * qR)
    rc, f = ob.face_flux("laxfriedrichs", qL, qR, n)
    assert rc == 0 and rel(f, want) < 1e-14
    _, fr = ob.rusanov(qL, qR, n)
    assert rel(f, fr) > 1e-3


def roe_textbook(qL, qR, n, g=GAMMA):
    """Roe's flux from explicit wave strengths of the Roe-averaged
    linearisation (Toro, 3rd ed., section 11.3, written for a unit normal),
    with Harten's fix |lambda| -> (lambda^2/(2 delta) + delta)/2 below
    2 delta, delta = a/10 -- and with the two liberties the reference takes
    (SF_roe.cpp:62-99), stated here explicitly:
      * it builds the eigenvectors at the state (rho~, rho~ u~, rho~ h~),
        whose energy slot holds rho~ h~, so their sound speed is
        c' = sqrt(gamma) a~ and their enthalpy H' = h~ + a~^2, while the
        wave speeds use a~;
      * it applies |u_n + a~| to the u_n - c' wave and |u_n - a~| to the
        u_n + c' wave."""
    def unpack(q):
        rho = q[0]
        vel = q[1:4] / rho
        p = (g - 1.0) * (q[4] - 0.5 * rho * vel @ vel)
        return rho, vel, p, (q[4] + p) / rho
    rL, vL, pL, hL = unpack(qL)
    rR, vR, pR, hR = unpack(qR)
    sL, sR = math.sqrt(rL), math.sqrt(rR)
    vt = (sL * vL + sR * vR) / (sL + sR)
    ht = (sL * hL + sR * hR) / (sL + sR)
    at = math.sqrt((g - 1.0) * (ht - 0.5 * vt @ vt))
    rt = sL * sR
    un = vt @ n
    cp = math.sqrt(g) * at
    hp = ht + at * at
    dp = pR - pL
    dun = (vR - vL) @ n
    drho = rR - rL
    dv = vR - vL
    # wave strengths and right eigenvectors
    a_minus = (dp - rt * cp * dun) / (2 * cp * cp)
    a_plus = (dp + rt * cp * dun) / (2 * cp * cp)
    a_ent = drho - dp / (cp * cp)
    r_minus = np.concatenate(([1.0], vt - cp * n, [hp - cp * un]))
    r_plus = np.concatenate(([1.0], vt + cp * n, [hp + cp * un]))
    r_ent = np.concatenate(([1.0], vt, [0.5 * vt @ vt]))
    dvt = dv - dun * n                    # tangential velocity jump
    r_shear = rt * np.concatenate(([0.0], dvt, [vt @ dvt]))

    def fix(lam):
        lam = abs(lam)
        d = 0.1 * at
        return 0.5 * (lam * lam / (2 * d) + d) if lam < 2 * d else lam
    dis = (fix(un + at) * a_minus * r_minus + fix(un - at) * a_plus * r_plus
           + fix(un) * (a_ent * r_ent + r_shear))

    def flux(rho, vel, p, q):
        vn = vel @ n
        return np.concatenate(([rho * vn], rho * vel * vn + p * n, [(q[4] + p) * vn]))
    return 0.5 * (flux(rL, vL, pL, qL) + flux(rR, vR, pR, qR) - dis)


def test_roe_matches_the_textbook_wave_decomposition():
    rng = np.random.default_rng(5)
    states = _states(rng, 40, speed=2.0)
    for qL, qR in zip(states[::2], states[1::2]):
        for n in NORMALS:
            rc, f = ob.face_flux("roe", qL, qR, n)
            assert rc == 0
            want = roe_textbook(qL, qR, n)
            assert np.max(np.abs(f - want)) < 1e-11 * max(1.0, np.max(np.abs(want))), (qL, qR, n)


def test_roe_swapped_acoustic_speeds_matter():
    # with the physical assignment the supersonic flux would be F(qL); the
    # reference's assignment is not, and the oracle follows the reference
    n = np.array([1.0, 0.0, 0.0])
    qL = prim(1.0, 3.0, 0.0, 0.0, 1.0)
    qR = prim(0.8, 2.8, 0.0, 0.0, 0.9)
    _, f = ob.face_flux("roe", qL, qR, n)
    _, fL = ob.physical_flux(qL, n)
    assert rel(f, fL) > 1e-4


@pytest.mark.parametrize("flux", ["steger", "laxfriedrichs", "roe"])
def test_splitting_fluxes_reject_nonphysical_states(flux):
    n = np.array([1.0, 0.0, 0.0])
    good = prim(1.0, 0.0, 0.0, 0.0, 1.0)
    bad = good.copy()
    bad[4] = -1.0
    assert ob.face_flux(flux, good, bad, n)[0] != 0
    assert ob.face_flux(flux, bad, good, n)[0] != 0


@pytest.mark.parametrize("scheme", ["weno3", "weno7"])
@pytest.mark.parametrize("flux", ["rusanov", "hllc", "steger", "laxfriedrichs", "roe"])
def test_constant_state_gives_the_physical_flux(scheme, flux):
    ns = ob.STENCIL[ob.SCHEMES[scheme]]
    q = prim(0.9, 0.3, -0.2, 0.1, 1.3)
    st = np.tile(q, (ns, 1))
    for axis in range(3):
        cof = np.zeros(3)
        cof[axis] = 2.5
        rc, f = ob.char_face_flux(st, cof, scheme, flux)
        _, fp = ob.physical_flux(q, cof / 2.5)
        assert rc == 0 and rel(f, 2.5 * fp) < 1e-12


# ---- the reference's frozen WENO7 Sod run ------------------------------------

def sod_weno7_state():
    """test/Sod/sodCase_weno7_t0p2: 101 x 21 nodes at unit spacing (z
    empty, dz = 0.1), leftHalf = nodes i < 50 at (1, 0, 1e5), the rest at
    (0.125, 0, 1e4), gamma 1.4."""
    nx, ny = 101, 21
    q = np.zeros((5, ny, nx))
    left = np.arange(nx) < nx // 2
    q[0] = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0e5, 1.0e4)
    q[4] = q[0] * (0.0 + p / ((GAMMA - 1.0) * q[0]))
    return q


def sod_weno7_desc():
    # the case names no rho / U / p condition for Left: that end carries none
    return ob.make_desc((101, 21, 1), (1, 1, 0), ("none_lo", "transmissive", "transmissive"),
                        (1.0, 1.0, 0.1), GAMMA, "weno7", "rusanov", "euler")


def sod_weno7_summary(q):
    """The quantities tools/check_sod_rusanov_regression.py takes from the
    final VTS: ranges of rho, p, |U| and trapezoidal integrals over the
    101 x 21 x 2 points (dz = 0.1, both z planes equal)."""
    q = q.reshape(5, 21, 101)
    rho = q[0]
    vel = q[1:4] / rho
    p = (GAMMA - 1.0) * (q[4] - 0.5 * rho * (vel ** 2).sum(axis=0))
    speed = np.sqrt((vel ** 2).sum(axis=0))
    wx = np.ones(101)
    wx[[0, -1]] = 0.5
    wy = np.ones(21)
    wy[[0, -1]] = 0.5
    W = np.outer(wy, wx) * 0.5 * 0.1 * 2
    return {
        "ranges": {"rho": [rho.min(), rho.max()], "p": [p.min(), p.max()],
                   "speed": [speed.min(), speed.max()]},
        "integrals": {"mass": (W * rho).sum(),
                      "momentum": [(W * q[1]).sum(), (W * q[2]).sum(), 0.0],
                      "energy": (W * q[4]).sum()},
    }


def test_oracle_reproduces_the_frozen_weno7_sod_baseline():
    base = json.loads(GOLDEN.read_text())
    q, dts = ob.run_driver(sod_weno7_desc(), sod_weno7_state(), 0.0, 0.2, 0.5,
                           write_time=0.05)
    assert len(dts) == base["steps"] == 454
    assert abs(dts.sum() - 0.2) < 1e-15
    # the log prints dt with 7 significant digits
    printed = np.array([float("%.7g" % x) for x in dts])
    ref = np.array(base["dt_sequence"])
    assert np.sum(printed == ref) >= 448
    assert np.max(np.abs(dts - ref) / ref) < 1e-6
    # one step is cut short to land on the output instant t = 0.05
    k = int(np.argmin(ref[:200]))
    assert ref[k] < 2e-5 and int(np.argmin(dts[:200])) == k
    assert abs(dts[:k + 1].sum() - 0.05) < 1e-15
    s = sod_weno7_summary(q)
    # the left-state region (rho max, p max, |U| min) sits away from the
    # forward-Euler oscillations at the right boundary: to ~1e-12
    for name, k in (("rho", 1), ("p", 1), ("speed", 0)):
        assert abs(s["ranges"][name][k] - base["ranges"][name][k]) \
            <= 1e-11 * abs(base["ranges"][name][k]), name
    # the extrema inside the oscillations move with last-bit rounding (the
    # reference's own FMA build moves them as much); integrals stay tight
    for name in ("rho", "p", "speed"):
        for k in (0, 1):
            assert abs(s["ranges"][name][k] - base["ranges"][name][k]) \
                <= 1e-6 * abs(base["ranges"][name][k]), name
    assert abs(s["integrals"]["mass"] - base["integrals"]["mass"]) < 1e-10 * base["integrals"]["mass"]
    assert abs(s["integrals"]["energy"] - base["integrals"]["energy"]) \
        < 1e-10 * base["integrals"]["energy"]
    mx = base["integrals"]["momentum"][0]
    assert abs(s["integrals"]["momentum"][0] - mx) < 1e-10 * mx
    assert abs(s["integrals"]["momentum"][1]) < 1e-9 * mx


def test_an_end_without_a_condition_is_not_a_zero_gradient_end():
    # declaring Left zeroGradient instead changes the left-region extrema
    # far beyond rounding: the baseline pins which one the case means
    base = json.loads(GOLDEN.read_text())
    d = sod_weno7_desc()
    d.bc[0] = ob.BCS["transmissive"]
    q, _ = ob.run_driver(d, sod_weno7_state(), 0.0, 0.2, 0.5, write_time=0.05)
    s = sod_weno7_summary(q)
    assert abs(s["ranges"]["rho"][1] - base["ranges"]["rho"][1]) > 1e-2


# ---- Lax-Wendroff -------------------------------------------------------------

def test_lax_wendroff_consistency_and_contact_known_answer():
    rng = np.random.default_rng(13)
    for q in _states(rng, 30):
        for n in NORMALS:
            rc, f = ob.lax_wendroff(q, q, n, 0.01, 0.1)
            _, fp = ob.physical_flux(q, n)
            assert rc == 0 and rel(f, fp) < 1e-12
    # a contact moving at u with uniform pressure: the Richtmyer half state
    # is the advected density rho* = (rhoL+rhoR)/2 - (lambda u / 2)(rhoR - rhoL),
    # lambda = dt / spacing, and the mass flux is u rho*
    u, p, dt, h = 0.7, 1.3, 0.02, 0.1
    qL = prim(1.0, u, 0.0, 0.0, p)
    qR = prim(0.4, u, 0.0, 0.0, p)
    lam = dt / h
    rho_star = 0.5 * (1.0 + 0.4) - 0.5 * lam * u * (0.4 - 1.0)
    rc, f = ob.lax_wendroff(qL, qR, np.array([1.0, 0.0, 0.0]), dt, h)
    assert rc == 0
    assert abs(f[0] - u * rho_star) < 1e-14
    assert abs(f[1] - (u * u * rho_star + p)) < 1e-13
    # it needs a positive dt and spacing (SF_lw.cpp:20-27)
    assert ob.lax_wendroff(qL, qR, np.array([1.0, 0.0, 0.0]), 0.0, h)[0] != 0
    assert ob.lax_wendroff(qL, qR, np.array([1.0, 0.0, 0.0]), dt, 0.0)[0] != 0
