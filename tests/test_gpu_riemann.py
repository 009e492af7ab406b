"""Point parity of the device face arithmetic (the functions the sweeps
use, through sf_riemann_flux / sf_face_flux) against the oracle on the same
states, and the HLLC wave-fan relations (tests/hllc_fan.py) on the device's
own HLLC.  Tolerances: 1e-12 of each face's flux scale for the face
fluxes, 1e-11 of it for the characteristic WENO face flux (the device forms
reciprocals and square roots from refined hardware estimates, within an ulp
of the correctly rounded values); the fan relations to 1e-9 of the flux
scale, as for the oracle."""
import numpy as np
import pytest

from oracle import bind as ob
from sonicsolver_amd._native import face_flux, riemann_flux
from tests import hllc_fan as hf

pytestmark = pytest.mark.gpu

FLUXES = ["rusanov", "hllc", "steger", "laxfriedrichs", "roe"]


def _random_pairs(count, seed, umax):
    rng = np.random.default_rng(seed)
    qL = np.array([hf.conserved(rng.uniform(0.1, 2.0), rng.uniform(-umax, umax, 3),
                                rng.uniform(0.1, 5.0)) for _ in range(count)])
    qR = np.array([hf.conserved(rng.uniform(0.1, 2.0), rng.uniform(-umax, umax, 3),
                                rng.uniform(0.1, 5.0)) for _ in range(count)])
    return qL, qR


@pytest.mark.parametrize("flux", FLUXES + ["laxwendroff"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_riemann_fluxes_match_the_oracle(flux, axis):
    print(f"case riemann {flux} axis {axis}", flush=True)
    n = hf.NORMALS[axis]
    for seed, umax in ((1, 0.5), (2, 3.0)):   # subsonic and supersonic mixes
        qL, qR = _random_pairs(400, seed + 10 * axis, umax)
        lw = 0.37
        dev, flags = riemann_flux(flux, axis, qL, qR, lw=lw)
        checked = 0
        for i in range(len(qL)):
            if flux == "laxwendroff":
                rc, ref = ob.lax_wendroff(qL[i], qR[i], n, lw * 0.01, 0.01)
            else:
                rc, ref = ob.face_flux(flux, qL[i], qR[i], n)
            # a pair the oracle rejects (the Lax-Wendroff half state of two
            # fast streams can lose positivity) is rejected on the device too
            assert (rc != 0) == (flags[i] != 0), (flux, axis, i, rc, flags[i])
            if rc != 0:
                continue
            checked += 1
            # relative to the face's flux scale: the physical fluxes of its
            # two states (a split flux can cancel to zero, as between two
            # states that leave the face supersonically)
            scale = max(np.abs(ref).max(), np.abs(hf.euler_flux(qL[i], n)).max(),
                        np.abs(hf.euler_flux(qR[i], n)).max())
            assert np.abs(dev[i] - ref).max() <= 1e-12 * scale, (flux, axis, i, dev[i], ref)
        assert checked >= len(qL) // 2


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_device_hllc_fan(axis):
    """The device HLLC satisfies the wave-fan relations that pin the
    oracle's: one call evaluates every ray of every pair."""
    n = hf.NORMALS[axis]
    pairs = hf.subsonic_pairs(6, seed=70 + axis)
    all_rays = [hf.fan_rays(qL, qR, n) for qL, qR in pairs]
    lefts, rights = [], []
    for (qL, qR), rays in zip(pairs, all_rays):
        a, b = hf.shifted_pairs(qL, qR, n, rays)
        lefts.append(a)
        rights.append(b)
    dev, flags = riemann_flux("hllc", axis, np.concatenate(lefts), np.concatenate(rights))
    assert not flags.any()
    k = 0
    for (qL, qR), rays in zip(pairs, all_rays):
        G = hf.fan_from_fluxes(rays, dev[k:k + len(rays)], n)
        k += len(rays)
        hf.check_fan(qL, qR, n, rays, G)


def _stencils(count, seed, jump):
    """Six-node stencils: smooth profiles plus, for `jump`, a discontinuity
    between random nodes."""
    rng = np.random.default_rng(seed)
    out = np.empty((count, 6, 5))
    x = np.arange(6) - 2.5
    for i in range(count):
        rho = 1.0 + 0.3 * np.sin(rng.uniform(0, 6) + rng.uniform(0.1, 1.0) * x)
        u = rng.uniform(-0.5, 0.5, 3)[None, :] + 0.1 * np.cos(x)[:, None] * rng.uniform(-1, 1, 3)
        p = 1.0 + 0.2 * np.cos(rng.uniform(0, 6) + 0.7 * x)
        if jump:
            cut = rng.integers(1, 6)
            rho[cut:] *= rng.uniform(0.1, 0.6)
            p[cut:] *= rng.uniform(0.1, 0.6)
        for s in range(6):
            out[i, s] = hf.conserved(rho[s], u[s], p[s])
    return out


@pytest.mark.parametrize("scheme", ["weno5", "teno5"])
@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_characteristic_face_flux_matches_the_oracle(scheme, flux, axis):
    print(f"case face {scheme}/{flux} axis {axis}", flush=True)
    area = 0.75
    cof = np.zeros(3)
    cof[axis] = area
    for jump in (False, True):
        st = _stencils(300, 5 + axis + 3 * jump, jump)
        dev, flags = face_flux(scheme, flux, axis, st, area=area)
        assert not flags.any()
        for i in range(len(st)):
            rc, ref = ob.char_face_flux(st[i], cof, scheme=scheme, flux=flux)
            assert rc == 0
            scale = max(np.abs(ref).max(), 1e-300)
            assert np.abs(dev[i] - ref).max() <= 1e-11 * scale, (scheme, flux, axis, i)


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_teno5_non_finite_gamma_sum_is_a_fault(flux):
    """A stencil whose first three nodes are equal (beta_0 = 0) next to a
    jump of 1e40 in density: tau5 / (beta_0 + 1e-40) overflows gamma_0, the
    sum of the gammas is not finite, and the reference stops
    (requireFiniteGammaSum, SF_TENO5.cpp:33-40).  The oracle reports it and
    so does the device."""
    print(f"case teno5 gamma overflow {flux}", flush=True)
    st = np.empty((1, 6, 5))
    for s in range(6):
        rho = 1.0 if s < 3 else 1.0e40
        st[0, s] = hf.conserved(rho, np.zeros(3), 1.0)
    cof = np.array([1.0, 0.0, 0.0])
    rc, _ = ob.char_face_flux(st[0], cof, scheme="teno5", flux=flux)
    assert rc != 0
    _, flags = face_flux("teno5", flux, 0, st, area=1.0)
    assert flags[0] != 0
