"""Properties of the HIP path at sizes the oracle cannot run: on a fully
periodic grid every face is evaluated by the same instructions wherever it
lies (tile, march segment, periodic copy), so translating the initial field
along an axis translates the result bit for bit.  The grid is large enough
that the y and z marches use their 64-node segments."""
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
