"""The provider-level solver on the device (recipe names -> compiled plan ->
op binding -> time driver) against the oracle's restatement of the
reference's time driver (itself bit-exact with the reference's own
Time::Driver, tests/test_oracle_vs_reference.py)."""
import numpy as np
import pytest

from oracle import bind as ob
from sonicsolver_amd import Solver, make_desc, workloads as wl
from tests.parity import assert_close

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("recipe,time", [("weno5Rusanov", "SSPRK3"), ("teno5HLLC", "classicalRK4"),
                                         ("weno5HLLC", "forwardEuler")])
def test_driver_to_end_time_with_outputs(recipe, time):
    print(f"provider {recipe}/{time}", flush=True)
    scheme = recipe[:5].lower()
    flux = "hllc" if recipe.lower().endswith("hllc") else "rusanov"
    tname = {"SSPRK3": "ssprk3", "classicalRK4": "rk4", "forwardEuler": "euler"}[time]
    kw, q = wl.sod_1d(300, scheme=scheme, flux=flux, time=tname)
    end, wt = 0.04, 0.0125
    with Solver(make_desc(**kw), recipe, time, cfl=0.4) as s:
        s.upload(q)
        steps, t, outputs = s.run(end, write_interval_time=wt)
        out = s.download()
        dts = s.dt_history()
    ref, rdts = ob.run_driver(ob.make_desc(**kw), q, 0.0, end, 0.4, write_time=wt)
    assert steps == len(rdts)
    np.testing.assert_allclose(dts, rdts, rtol=1e-12, atol=0)
    assert abs(t - end) < 1e-14
    assert [round(o[1] / wt) for o in outputs] == [1, 2, 3]
    assert_close(out, ref, what=f"provider {recipe}/{time}")


def test_driver_end_step_and_step_outputs():
    kw, q = wl.vortex_2d(48)
    with Solver(make_desc(**kw), "teno5HLLC", "SSPRK3", cfl=0.5) as s:
        s.upload(q)
        steps, t, outputs = s.run(100.0, end_step=6, write_by_step=True, write_interval_steps=2)
        out = s.download()
    ref, rdts = ob.run_driver(ob.make_desc(**kw), q, 0.0, 100.0, 0.5, end_step=6,
                              write_by_step=True, write_steps=2)
    assert steps == 6 and [o[0] for o in outputs] == [2, 4, 6]
    assert abs(t - rdts.sum()) < 1e-13
    assert_close(out, ref, what="provider vortex")
