"""The wave fan of an HLLC flux, recovered from the flux alone, and the
relations that the mathematics fixes for it (Toro, "Riemann Solvers and
Numerical Methods for Fluid Dynamics", 3rd ed., section 10.4).

An approximate Riemann solver gives, for every ray x/t = w, a state U(w)
and the flux across a surface moving with speed w along the normal,
G(w) = F - w U.  Seen from a frame moving with w, the face sits at x/t = 0
of the shifted problem, so G(w) follows from one flux evaluation on the
shifted states (Galilean transform below).  HLLC's fan has four constant
states U_L | U*_L | U*_R | U_R separated at S_L < S* < S_R, so G(w) is
continuous and piecewise linear with slopes -U_L, -U*_L, -U*_R, -U_R.  The
checks, none of which uses how the flux forms its wave speeds or star
states:

  * the outer slopes are the data (upwinding beyond the fastest waves);
  * across the contact the mass flux vanishes, the tangential momentum flux
    vanishes and the energy flux is the pressure work: G(S*) =
    (0, p* n, p* S*) with one p* from both star states (u*_L = u*_R = S*,
    p*_L = p*_R);
  * G is continuous at S_L, S* and S_R (the Rankine-Hugoniot conditions
    across each wave, F*_R = F*_L + S* (U*_R - U*_L) at the contact), which
    with the constant states is the integral condition
    (S* - S_L) U*_L + (S_R - S*) U*_R = S_R U_R - S_L U_L - (F_R - F_L);
  * each star state carries its side's tangential velocity and moves with
    the contact: m*_n / rho* = S*.

The energy of the star states enters G only through S_K (E*_K - E_K), so
a star energy that is off by any amount moves G_E(S*) away from p* S*."""
import numpy as np


def conserved(rho, u, p, gamma=1.4):
    u = np.asarray(u, dtype=float)
    return np.array([rho, *(rho * u), p / (gamma - 1.0) + 0.5 * rho * u @ u])


def euler_flux(q, n, gamma=1.4):
    """The Euler flux F(q) . n."""
    u = q[1:4] / q[0]
    p = (gamma - 1.0) * (q[4] - 0.5 * q[0] * u @ u)
    un = u @ n
    return np.array([q[0] * un, *(q[1:4] * un + p * n), (q[4] + p) * un])


def shift(q, w, n):
    """The conserved state seen from a frame moving with velocity w n."""
    rho, m, E = q[0], q[1:4], q[4]
    mn = m @ n
    return np.array([rho, *(m - rho * w * n), E - w * mn + 0.5 * rho * w * w])


def lab_flux(fp, w, n):
    """G(w) = F - w U in the lab frame from the flux fp computed in the
    frame moving with w n (exact for the Euler equations)."""
    g = np.empty(5)
    g[0] = fp[0]
    g[1:4] = fp[1:4] + w * n * fp[0]
    g[4] = fp[4] + w * (fp[1:4] @ n) + 0.5 * w * w * fp[0]
    return g


def fan_rays(qL, qR, n, gamma=1.4, count=6001):
    """Rays spanning the fan of (qL, qR) with a margin on both sides."""
    speeds = []
    for q in (qL, qR):
        u = q[1:4] / q[0]
        p = (gamma - 1.0) * (q[4] - 0.5 * q[0] * u @ u)
        c = np.sqrt(gamma * p / q[0])
        speeds += [u @ n - c, u @ n + c]
    span = max(speeds) - min(speeds)
    return np.linspace(min(speeds) - 0.5 * span, max(speeds) + 0.5 * span, count)


def shifted_pairs(qL, qR, n, rays):
    """(left, right) states of the problem seen from each ray's frame."""
    return (np.array([shift(qL, w, n) for w in rays]),
            np.array([shift(qR, w, n) for w in rays]))


def fan_from_fluxes(rays, fluxes, n):
    """G on the rays from the fluxes of the shifted problems."""
    return np.array([lab_flux(f, w, n) for w, f in zip(rays, fluxes)])


def _runs(rays, G, rtol):
    """Maximal runs of intervals with one slope: [(first, last, U)]."""
    dw = np.diff(rays)
    slopes = -np.diff(G, axis=0) / dw[:, None]
    scale = np.abs(slopes).max(axis=0) + 1e-300
    runs = []
    start = 0
    for i in range(1, len(slopes) + 1):
        if i == len(slopes) or np.any(np.abs(slopes[i] - slopes[start]) > rtol * scale):
            runs.append((start, i - 1))
            start = i
    out = []
    for a, b in runs:
        if b - a + 1 >= 4:   # drop the single intervals that straddle a kink
            mid = (a + b) // 2
            out.append((a, b, slopes[mid]))
    return out


def check_fan(qL, qR, n, rays, G, tol=1e-9, gamma=1.4):
    """Assert the HLLC fan relations on G(rays); returns (S_L, S*, S_R, p*)."""
    runs = _runs(rays, G, 1e-7)
    assert len(runs) == 4, f"expected four constant states in the fan, found {len(runs)}"
    (a0, b0, UL), (a1, b1, UsL), (a2, b2, UsR), (a3, b3, UR) = runs
    scale = np.abs(np.concatenate([qL, qR])).max()
    assert np.abs(UL - qL).max() <= tol * scale, "left slope is not the left state"
    assert np.abs(UR - qR).max() <= tol * scale, "right slope is not the right state"
    # the line of each run: G(w) = A - w U
    A = [G[b] + rays[b] * U for (_, b, U) in runs]
    Us = [UL, UsL, UsR, UR]

    def meet(k):
        # where lines k and k+1 meet, from the density component
        return (A[k][0] - A[k + 1][0]) / (Us[k][0] - Us[k + 1][0])

    SL, Ss, SR = meet(0), meet(1), meet(2)
    assert rays[b0] <= SL <= rays[a1 + 1] and rays[b1] <= Ss <= rays[a2 + 1]
    assert rays[b2] <= SR <= rays[a3 + 1]
    gscale = np.abs(G).max()
    for k, w in enumerate((SL, Ss, SR)):
        # continuity of every component where the lines meet (Rankine-Hugoniot)
        gl = A[k] - w * Us[k]
        gr = A[k + 1] - w * Us[k + 1]
        assert np.abs(gl - gr).max() <= tol * gscale, f"G jumps at wave {k}"
    # the outer lines at w = 0 are the physical fluxes of the data
    FL, FR = A[0], A[3]
    for F, q in ((FL, qL), (FR, qR)):
        assert np.abs(F - euler_flux(q, n, gamma)).max() <= tol * gscale
    # integral condition over the fan
    lhs = (Ss - SL) * UsL + (SR - Ss) * UsR
    rhs = SR * qR - SL * qL - (FR - FL)
    assert np.abs(lhs - rhs).max() <= tol * max(np.abs(rhs).max(), gscale)
    # the contact: no mass or tangential momentum through it, pressure work
    Gc = A[1] - Ss * UsL
    pstar = Gc[1:4] @ n
    t = Gc[1:4] - pstar * n
    assert abs(Gc[0]) <= tol * gscale, "mass flows through the contact"
    assert np.abs(t).max() <= tol * gscale, "tangential momentum flows through the contact"
    assert abs(Gc[4] - pstar * Ss) <= tol * gscale, \
        f"energy flux at the contact {Gc[4]:.17g} is not the pressure work {pstar * Ss:.17g}"
    for U, q in ((UsL, qL), (UsR, qR)):
        un = (U[1:4] @ n) / U[0]
        assert abs(un - Ss) <= tol * max(1.0, abs(Ss)), "a star state does not move with S*"
        ut = U[1:4] / U[0] - un * n
        qt = q[1:4] / q[0] - (q[1:4] @ n / q[0]) * n
        assert np.abs(ut - qt).max() <= tol * max(1.0, np.abs(qt).max()), \
            "a star state does not carry its side's tangential velocity"
    return SL, Ss, SR, pstar


def subsonic_pairs(count, seed):
    """Random subsonic pairs whose contact does not move with either side's
    velocity (S* != u_K), as conserved states."""
    rng = np.random.default_rng(seed)
    pairs = []
    while len(pairs) < count:
        rl, rr = rng.uniform(0.2, 2.0, 2)
        pl, pr = rng.uniform(0.2, 3.0, 2)
        ul = rng.uniform(-0.4, 0.4, 3)
        ur = rng.uniform(-0.4, 0.4, 3)
        if abs(pl - pr) < 0.2 and abs(ul[0] - ur[0]) < 0.1:
            continue   # keep the contact well away from u_L and u_R
        pairs.append((conserved(rl, ul, pl), conserved(rr, ur, pr)))
    return pairs


NORMALS = [np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0])]
