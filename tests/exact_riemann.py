"""Exact solution of the 1D Riemann problem for a perfect gas (Toro,
"Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4:
pressure function (4.6)-(4.8), Newton iteration (4.44), sampling (4.4)).
Used as an accuracy check of the oracle, independent of the reference."""
import math

import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        s = math.sqrt(A / (p + B))
        return (p - pk) * s, s * (1.0 - 0.5 * (p - pk) / (B + p))
    pr = p / pk  # rarefaction
    return (2.0 * ck / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0),
            1.0 / (rho * ck) * pr ** (-(g + 1.0) / (2.0 * g)))


def star_state(rl, ul, pl, rr, ur, pr, g=1.4):
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dl = _f(p, rl, pl, cl, g)
        fr, dr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dl + dr)
        pn = max(1e-12, p - dp)
        if abs(pn - p) <= 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def sample(x, t, x0, left, right, g=1.4):
    """Primitive (rho, u, p) at positions x and time t > 0."""
    rl, ul, pl = left
    rr, ur, pr = right
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    out = np.zeros((3, len(x)))
    gm = (g - 1.0) / (g + 1.0)
    for n, xi in enumerate(np.asarray(x)):
        s = (xi - x0) / t
        if s <= us:
            if ps > pl:
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sl:
                    out[:, n] = (rl, ul, pl)
                else:
                    out[:, n] = (rl * (ps / pl + gm) / (gm * ps / pl + 1.0), us, ps)
            else:
                shl = ul - cl
                cml = cl * (ps / pl) ** ((g - 1) / (2 * g))
                stl = us - cml
                if s <= shl:
                    out[:, n] = (rl, ul, pl)
                elif s > stl:
                    out[:, n] = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    rho = rl * (c / cl) ** (2 / (g - 1))
                    out[:, n] = (rho, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sr:
                    out[:, n] = (rr, ur, pr)
                else:
                    out[:, n] = (rr * (ps / pr + gm) / (gm * ps / pr + 1.0), us, ps)
            else:
                shr = ur + cr
                cmr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                str_ = us + cmr
                if s >= shr:
                    out[:, n] = (rr, ur, pr)
                elif s <= str_:
                    out[:, n] = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    rho = rr * (c / cr) ** (2 / (g - 1))
                    out[:, n] = (rho, u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out
