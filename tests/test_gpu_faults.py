"""Fail-fast on the device, as the reference stops at the first
non-physical stage (validateEquationState, SF_equationState.h:25-57): once a
fault flag is raised every later kernel of the grid returns at entry, the
run reports the faulting step, and the grid keeps the state at the start of
that step.  The faults here are real numerical failures (two rarefactions
strong enough to open a vacuum), not planted ones: each kernel still exits
normally.  The oracle (the C restatement) fails at the same step, and the
state the device keeps equals the oracle's state after the steps before it
(relative L1 / Linf <= 1e-11)."""
import numpy as np
import pytest

from oracle import bind as ob
from sonicsolver_amd import Grid, Solver, SonicError, make_desc, workloads as wl
from sonicsolver_amd.batch import Batch, item_nodes, make_item
from tests.parity import assert_close

pytestmark = pytest.mark.gpu

# (u of each side, pressure, scheme, flux, time): the oracle fails within
# the first ten steps of each, on a non-finite flux or a non-physical state
VACUUM_CASES = [(3.0, 1.0, "weno5", "hllc", "ssprk3"), (2.0, 0.4, "weno5", "rusanov", "ssprk3"),
                (3.0, 1.0, "teno5", "hllc", "ssprk3"), (2.0, 0.4, "weno5", "hllc", "rk4")]


def _vacuum(u, p, scheme, flux, time, cells=1200):
    kw, q = wl.riemann_1d(cells, (1.0, -u, p), (1.0, u, p), 0.5)
    desc_kw = dict(n=(kw["n"][0], 1, 1), active=(1, 0, 0), d=(kw["d"][0], 1.0, 1.0),
                   scheme=scheme, flux=flux, time=time)
    return desc_kw, q


def _oracle_failure(desc_kw, q, cfl, limit=100):
    """(failing step, state after the steps before it, their dt values)."""
    desc = ob.make_desc(**desc_kw)
    state, dts = q.copy(), []
    for s in range(limit):
        try:
            state, dt = ob.run(desc, state, 1, cfl)
        except ob.OracleError:
            return s + 1, state, np.array(dts)
        dts.append(dt[0])
    raise AssertionError("the oracle did not fail")


@pytest.mark.parametrize("u,p,scheme,flux,time", VACUUM_CASES)
def test_run_stops_at_the_oracles_failing_step(u, p, scheme, flux, time):
    desc_kw, q = _vacuum(u, p, scheme, flux, time)
    print(f"vacuum u={u} p={p} {scheme}/{flux}/{time}", flush=True)
    fail_step, ref, ref_dts = _oracle_failure(desc_kw, q, 0.5)
    with Grid(make_desc(**desc_kw)) as g:
        g.upload(q)
        dts = np.zeros(40)
        with pytest.raises(SonicError) as e:
            dts[:] = g.run(40, 0.5)
        assert e.value.code == -3
        assert f"in step {fail_step}," in str(e.value), str(e.value)
        flags, step, phase = g.fault()
        assert flags != 0 and step == fail_step and 0 <= phase <= 7
        # the grid holds the state at the start of the faulting step, and
        # its clock counts the steps before it
        t, steps = g.clock()
        assert steps == fail_step - 1
        assert abs(t - ref_dts.sum()) <= 1e-12 * ref_dts.sum()
        out = g.download()
        assert_close(out, ref, what=f"state before the faulting step {fail_step}")
        # the fault record is kept until cleared: a further run does nothing
        with pytest.raises(SonicError):
            g.run(3, 0.5)
        assert g.clock()[1] == fail_step - 1
        np.testing.assert_array_equal(g.download(), out)
        # clearing the flags clears the record
        g.flags(clear=True)
        assert g.fault() == (0, 0, -1)


def test_op_by_op_stepping_restores_on_check():
    desc_kw, q = _vacuum(3.0, 1.0, "weno5", "hllc", "ssprk3")
    fail_step, ref, _ = _oracle_failure(desc_kw, q, 0.5)

    def step(g, dt=None):
        g.prepare()
        if dt is None:
            g.compute_dt(0.5)
        else:
            g.set_dt(dt)
        g.begin()
        for s in range(g.stage_count()):
            g.stage(s)
        g.commit()

    # steps queued after the fault (with a given dt, so nothing checks in
    # between) do nothing; the check finds the faulting step
    with Grid(make_desc(**desc_kw)) as g:
        g.upload(q)
        for k in range(fail_step + 3):
            step(g, None if k < fail_step else 1e-3)
        flags, at, phase = g.fault()
        assert flags != 0 and at == fail_step and 0 <= phase <= 7
        assert g.clock()[1] == fail_step - 1
        assert_close(g.download(), ref, what="op-by-op state before the fault")
    # compute_dt of the step after the faulting one reports it
    with Grid(make_desc(**desc_kw)) as g:
        g.upload(q)
        with pytest.raises(SonicError) as e:
            for _ in range(fail_step + 1):
                step(g)
        assert f"in step {fail_step}," in str(e.value)
        assert g.clock()[1] == fail_step - 1
        assert_close(g.download(), ref, what="op-by-op state before the fault")


def test_provider_run_reports_the_faulting_step():
    desc_kw, q = _vacuum(2.0, 0.4, "weno5", "rusanov", "ssprk3")
    fail_step, ref, ref_dts = _oracle_failure(desc_kw, q, 0.5)
    with Solver(make_desc(**desc_kw), "weno5Rusanov", "SSPRK3", cfl=0.5) as s:
        s.upload(q)
        with pytest.raises(SonicError) as e:
            s.run(10.0)
        assert f"in step {fail_step}," in str(e.value), str(e.value)
        np.testing.assert_allclose(s.dt_history(), ref_dts, rtol=1e-12, atol=0)
        assert_close(s.download(), ref, what="provider state before the fault")


def test_batch_fault_in_mid_step_with_multiwave_tiles():
    """A multi-tile problem that fails in the middle of the run, under
    256-lane x tiles (four waves meeting at barriers): every wave of a tile
    takes the same liveness decision, the failing problem is counted up to
    the step before its fault, and the others are bit-identical to their
    runs without it."""
    probs = []
    for i, kw, q in wl.batch_riemann_1d(6, seed=31, lo=300, hi=2000):
        probs.append((kw, q))
    kwb, qb = wl.riemann_1d(1200, (1.0, -3.0, 1.0), (1.0, 3.0, 1.0), 0.5)
    kwb.update(scheme="weno5", flux="hllc", cfl=0.5)
    desc_b = dict(n=(kwb["n"][0], 1, 1), active=(1, 0, 0), d=(kwb["d"][0], 1.0, 1.0),
                  scheme="weno5", flux="hllc", time="ssprk3")
    fail_step, ref_b, _ = _oracle_failure(desc_b, qb, 0.5)
    steps = fail_step + 6
    cap = sum(item_nodes(make_item(**kw)) for kw, _ in probs) + item_nodes(make_item(**kwb)) + 64

    def run(with_bad, graph):
        with Batch(cap, len(probs) + 1) as b:
            b.set_tiles(255, 32)
            b.set_schedule(False)
            bad = b.add(make_item(**kwb), qb) if with_bad else None
            slots = [b.add(make_item(**kw), q) for kw, q in probs]
            if with_bad:
                with pytest.raises(SonicError) as e:
                    b.step(steps, graph=graph)
                assert e.value.code == -3
                assert b.faults()[bad] != 0
                assert b.clock(bad)[1] == fail_step - 1
            else:
                b.step(steps, graph=graph)
            return [b.download(s) for s in slots], [b.clock(s) for s in slots]

    base, bclocks = run(False, True)
    for graph in (True, False):
        outs, clocks = run(True, graph)
        for k in range(len(probs)):
            assert np.array_equal(outs[k], base[k]), k
            assert clocks[k] == bclocks[k]
