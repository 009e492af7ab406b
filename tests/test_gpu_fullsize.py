"""The headline workload at its full size: the 512^3 periodic Taylor-Green
field with WENO5-JS, HLLC and SSPRK3 at CFL 0.4, as the benchmark runs it.
The oracle needs tens of minutes per step at this size, so the run is held
to properties that do not depend on the size: every conserved sum stays
constant to rounding (the flux-difference form telescopes on a periodic
grid), every node stays physical, the CFL steps are positive and finite,
and translating the field along z by a distance that crosses march
segments and x-sweep tiles translates the result bit for bit."""
import numpy as np
import pytest

from sonicsolver_amd import Grid, make_desc, workloads as wl

pytestmark = pytest.mark.gpu

N = 512
STEPS = 2
SHIFT = 129   # z planes: not a multiple of the 64-node march segment


@pytest.fixture(scope="module")
def tgv():
    return wl.tgv_3d(N)


def _run(kw, q):
    with Grid(make_desc(**kw)) as g:
        g.upload(q)
        dts = g.run(STEPS, 0.4)
        assert g.flags() == 0
        return g.download(), dts


def test_tgv_512_conserves_and_translates(tgv):
    kw, q = tgv
    print("case tgv 512^3 weno5 hllc ssprk3, 2 steps", flush=True)
    out, dts = _run(kw, q)
    assert len(dts) == STEPS and np.all(np.isfinite(dts)) and np.all(dts > 0.0)
    assert np.all(out[0] > 0.0)
    for v in range(5):
        # relative to the magnitude of the variable (z momentum starts at 0)
        scale = max(np.abs(q[v]).sum(), np.abs(out[v]).sum())
        assert abs(out[v].sum() - q[v].sum()) <= 1e-12 * scale, v
    # pressure stays positive: E - |m|^2 / (2 rho) > 0
    ke = 0.5 * (out[1] ** 2 + out[2] ** 2 + out[3] ** 2) / out[0]
    assert np.all(out[4] - ke > 0.0)
    del ke

    print("case tgv 512^3 translated by", SHIFT, "z planes", flush=True)
    shape = (5, N, N, N)   # (var, z, y, x)
    moved = np.roll(q.reshape(shape), SHIFT, axis=1).reshape(5, -1)
    out_m, dts_m = _run(kw, moved)
    del moved
    assert np.array_equal(dts, dts_m)
    back = np.roll(out_m.reshape(shape), -SHIFT, axis=1).reshape(5, -1)
    assert np.array_equal(out, back)
