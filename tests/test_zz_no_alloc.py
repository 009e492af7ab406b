"""sf_run_static, the entry point that allocates nothing, against sf_run
(the grid path it restates) through the C ABI: the same kernels on the same
inputs give the same state and dt sequence bit for bit, on two calls in a
row with the second smaller than the first (the shared device array keeps
the first call's contents), on a run that faults (the state at the start
of the faulting step comes back), and a grid beyond the array's size is
rejected before any work.  The inputs and outputs are device arrays the
test allocates, as a caller would."""
import numpy as np
import pytest

from sonicsolver_amd import Grid, SonicError, make_desc, workloads as wl
from sonicsolver_amd._native import run_static
from tests.devmem import DeviceArray

pytestmark = pytest.mark.gpu


def _grid_run(kw, q, nsteps, cfl):
    with Grid(make_desc(**kw)) as g:
        g.upload(q)
        try:
            dts = g.run(nsteps, cfl)
            err = None
        except SonicError as e:
            dts, err = None, e
        return g.download(), dts, err


def _static_run(kw, q, nsteps, cfl):
    a = np.ascontiguousarray(q, dtype=np.float64).reshape(5, -1)
    with DeviceArray(a.size) as dev, DeviceArray(max(nsteps, 1)) as dts:
        dev.upload(a)
        dts.upload(np.full(max(nsteps, 1), -1.0))
        err = None
        try:
            need = run_static(make_desc(**kw), dev.ptr, nsteps, cfl, 1e300, dts.ptr, 0)
            assert need > 0
        except SonicError as e:
            err = e
        return dev.download().reshape(5, -1), dts.download()[:nsteps], err


def test_two_calls_in_a_row_equal_sf_run():
    cases = [(wl.vortex_2d(96), 12, 0.5),                                   # 2D TENO5/HLLC SSPRK3
             (wl.sod_1d(200, scheme="weno5", flux="rusanov", time="rk4"), 20, 0.4)]
    for (kw, q), nsteps, cfl in cases:
        print(f"case static n={kw['n']} {kw['scheme']}/{kw['flux']}/{kw['time']} {nsteps} steps",
              flush=True)
        ref, ref_dts, err = _grid_run(kw, q, nsteps, cfl)
        assert err is None
        out, dts, err = _static_run(kw, q, nsteps, cfl)
        assert err is None
        assert np.array_equal(out, ref)
        assert np.array_equal(dts, ref_dts)


def test_fault_returns_the_state_before_the_faulting_step():
    kw, q = wl.riemann_1d(1200, (1.0, -3.0, 1.0), (1.0, 3.0, 1.0), 0.5)
    desc_kw = dict(n=(kw["n"][0], 1, 1), active=(1, 0, 0), d=(kw["d"][0], 1.0, 1.0),
                   scheme="weno5", flux="hllc", time="ssprk3")
    print("case static vacuum", flush=True)
    ref, _, rerr = _grid_run(desc_kw, q, 40, 0.5)
    out, _, err = _static_run(desc_kw, q, 40, 0.5)
    assert rerr is not None and err is not None and err.code == -3
    step = str(rerr).split("in step ")[1].split(",")[0]
    assert f"in step {step}," in str(err)
    assert np.array_equal(out, ref)


def test_grid_beyond_the_array_is_rejected():
    kw = dict(n=(128, 128, 128), active=(1, 1, 1), bc=("periodic",) * 3, d=(0.1, 0.1, 0.1),
              scheme="weno5", flux="hllc", time="ssprk3")
    q = np.ones((5, 128 ** 3))
    q[4] = 2.5
    print("case static oversized", flush=True)
    out, _, err = _static_run(kw, q, 1, 0.4)
    assert err is not None and err.code == -1 and "bytes of device scratch" in str(err)
    assert np.array_equal(out, q)   # nothing ran
