"""Properties of the HIP path at sizes the oracle cannot run: on a fully
periodic grid every face is evaluated by the same instructions wherever it
lies (tile, march segment, periodic copy), so translating the initial field
along an axis translates the result bit for bit; and the flux-difference
form conserves every variable's sum.  The grid is large enough that the y
and z marches use their 64-node segments."""
import numpy as np
import pytest

from sonicsolver_amd import Grid, make_desc, workloads as wl

pytestmark = pytest.mark.gpu


N = (512, 512, 128)


@pytest.fixture(scope="module")
def field():
    return wl.smooth_field(N, seed=29).reshape(5, N[2], N[1], N[0])


@pytest.mark.parametrize("axis,shift", [(2, 37), (1, 101), (0, 5)])
def test_translation_is_exact_on_a_periodic_grid(field, axis, shift):
    n, q = N, field
    kw = dict(n=n, active=(1, 1, 1), bc=("periodic",) * 3, d=(0.1, 0.12, 0.09),
              scheme="weno5", flux="hllc", time="ssprk3")
    np_axis = 3 - axis   # (var, z, y, x)
    print("case translation axis", axis, "shift", shift, flush=True)
    outs = []
    for s in (0, shift):
        with Grid(make_desc(**kw)) as g:
            g.upload(np.roll(q, s, axis=np_axis).reshape(5, -1))
            dts = g.run(2, 0.4)
            outs.append((np.roll(g.download().reshape(q.shape), -s, axis=np_axis), dts))
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][0], outs[1][0])


def test_conservation_on_a_periodic_grid(field):
    """The flux-difference form is conservative: on a periodic grid the sum
    of every conserved variable is constant to rounding (relative to the sum
    of its magnitudes), for every time recipe."""
    q = field.reshape(5, -1)
    for time in ("euler", "ssprk3", "rk4"):
        kw = dict(n=N, active=(1, 1, 1), bc=("periodic",) * 3, d=(0.1, 0.12, 0.09),
                  scheme="weno5", flux="hllc", time=time)
        print("case conservation", time, flush=True)
        with Grid(make_desc(**kw)) as g:
            g.upload(q)
            g.run(3, 0.4)
            out = g.download()
        for v in range(5):
            scale = np.abs(q[v]).sum()
            assert abs(out[v].sum() - q[v].sum()) <= 1e-12 * scale, (time, v)
