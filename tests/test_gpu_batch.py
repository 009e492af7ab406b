"""Batched solves on the device against the oracle, and the properties the
batch promises: each problem's result does not depend on which problems
share the batch (bit for bit), nor on the launch schedule, tile shape,
persistent grid size or graph replay; admission by arena footprint; step
limits.  Tolerance against the oracle: relative L1 and Linf <= 1e-11."""
import numpy as np
import pytest

from oracle import bind as ob
from sonicsolver_amd import Grid, SonicError, make_desc, workloads as wl
from sonicsolver_amd.batch import Batch, make_item, item_nodes
from tests.parity import assert_close, rel_errors

pytestmark = pytest.mark.gpu

QUADS = [(0.5, 0.1, -0.2, 1.0), (1.0, -0.3, 0.2, 0.4), (0.2, 0.5, 0.5, 0.3), (1.5, 0.0, 0.1, 1.7)]


def _oracle(kw, q, steps):
    n = kw["n"]
    two_d = len(n) > 1 and n[1] > 1
    desc = ob.make_desc(n=(n[0], n[1] if two_d else 1, 1), active=(1, 1 if two_d else 0, 0),
                        d=(kw["d"][0], kw["d"][1] if two_d else 1.0, 1.0),
                        scheme=kw["scheme"], flux=kw["flux"], time="ssprk3")
    return ob.run(desc, q, steps, kw["cfl"])


def _mixed_problems(count=12, seed=7, lo=40, hi=700):
    probs = [(kw, q) for _, kw, q in wl.batch_riemann_1d(count, seed=seed, lo=lo, hi=hi)]
    for n, scheme, flux in [(9, "teno5", "hllc"), (33, "weno5", "hllc"), (70, "weno5", "rusanov"),
                            (129, "teno5", "rusanov")]:
        kw, q = wl.quadrant_2d(n, QUADS)
        kw.update(scheme=scheme, flux=flux, cfl=0.5)
        probs.append((kw, q))
    return probs


def _run(probs, steps, graph=True, fused=True, tiles=(127, 32), occupancy=8, capacity=None):
    cap = capacity or sum(item_nodes(make_item(**kw)) for kw, _ in probs) + 64
    with Batch(cap, len(probs) + 1) as b:
        b.set_schedule(fused)
        b.set_tiles(*tiles)
        b.set_occupancy(occupancy)
        slots = [b.add(make_item(**kw), q) for kw, q in probs]
        b.step(steps, graph=graph)
        outs = [b.download(s) for s in slots]
        clocks = [b.clock(s) for s in slots]
    return outs, clocks


def test_batch_against_oracle():
    probs = _mixed_problems()
    steps = 25
    print(f"batch of {len(probs)} problems, {steps} steps", flush=True)
    outs, clocks = _run(probs, steps)
    for (kw, q), out, (t, k) in zip(probs, outs, clocks):
        ref, dts = _oracle(kw, q, steps)
        assert k == steps
        assert abs(t - dts.sum()) <= 1e-12 * dts.sum()
        assert_close(out, ref, what=f"batch problem n={kw['n']} {kw['scheme']}/{kw['flux']}")


def test_batch_riemann_full_length_range():
    """Lengths 256..8192 cells as in the batched configuration.  WENO5
    problems meet the bar after 60 steps, TENO5 problems after 20.

    TENO5 (cutoff 1e-5, epsilon 1e-40) decides its stencils on smoothness
    indicators that are rounding noise in near-flat regions, so on this data
    it amplifies last-bit differences chaotically: the reference's own
    sources built with FMA contraction (its release flags) leave their
    build without contraction by up to ~1e-6 (Linf) within 60 steps
    (tests/test_oracle_vs_reference.py), and the same device source run on
    the CPU with and without contraction leaves the oracle by 1e-7..1e-8.
    After 60 steps TENO5 problems are therefore held to what chaos cannot
    move: the scheme is conservative and the waves have not reached the
    ends, so over the run's own time t every conserved total changes by
    exactly -t/dx (F(q_right) - F(q_left)) of the two constant end states
    (checked to rounding for the device and the oracle alike); and to the
    same solution in the mean (relative L1 <= 1e-4)."""
    probs = [(kw, q) for _, kw, q in wl.batch_riemann_1d(8, seed=4096)]
    kw, q = wl.riemann_1d(8192, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.5)
    kw.update(scheme="teno5", flux="hllc", cfl=0.5)
    probs.append((kw, q))

    def flux_x(qv):
        rho, u, _, _, p = wl.primitive(qv.reshape(5, 1))
        return np.array([rho[0] * u[0], 0.0, 0.0, 0.0, 0.0]) + np.array(
            [0.0, rho[0] * u[0] * u[0] + p[0], 0.0, 0.0, (qv[4] + p[0]) * u[0]])

    def conserved_within(qfinal, q0, t, dx, what):
        expected = -t / dx * (flux_x(q0[:, -1]) - flux_x(q0[:, 0]))
        for v in (0, 1, 4):
            scale = np.abs(q0[v]).sum() + abs(expected[v])
            change = qfinal[v].sum() - q0[v].sum()
            assert abs(change - expected[v]) <= 1e-11 * scale, (what, v)

    for steps in (20, 60):
        outs, clocks = _run(probs, steps)
        for (kw, q), out, (t, _) in zip(probs, outs, clocks):
            what = f"riemann 1d n={kw['n'][0]} {kw['scheme']}/{kw['flux']} {steps} steps"
            ref, dts = _oracle(kw, q, steps)
            if kw["scheme"] == "weno5" or steps == 20:
                assert_close(out, ref, what=what)
                continue
            l1, li = rel_errors(out, ref)
            print(f"{what}: device vs oracle rel L1 {l1.max():.2e} Linf {li.max():.2e}",
                  flush=True)
            q0 = np.asarray(q).reshape(5, -1)
            conserved_within(out, q0, t, kw["d"][0], what + " (device)")
            conserved_within(ref, q0, dts.sum(), kw["d"][0], what + " (oracle)")
            assert l1.max() <= 1e-4, what


def _short_problems(scheme, count):
    """The shortest `count` problems of one reconstruction among the first 24
    of the batched configuration's generator (lengths 256..1500 cells)."""
    probs = [(kw, q) for _, kw, q in wl.batch_riemann_1d(24, seed=4096, lo=256, hi=1500)
             if kw["scheme"] == scheme]
    probs.sort(key=lambda p: p[0]["n"][0])
    return probs[:count]


def test_batch_weno5_at_the_stated_300_steps():
    """WENO5 problems of the batched configuration (Rusanov and HLLC) for the
    full 300 steps: the reference built with and without FMA contraction
    agrees to 6e-14 there, and the device is held to 1e-11."""
    probs = _short_problems("weno5", 6)
    print(f"batch weno5 x{len(probs)}, 300 steps", flush=True)
    outs, clocks = _run(probs, 300)
    for (kw, q), out, (t, k) in zip(probs, outs, clocks):
        ref, dts = _oracle(kw, q, 300)
        assert k == 300 and abs(t - dts.sum()) <= 1e-12 * dts.sum()
        assert_close(out, ref, what=f"weno5/{kw['flux']} n={kw['n'][0]}, 300 steps")


def test_batch_teno5_at_300_steps_within_the_references_own_spread():
    """TENO5 problems of the batched configuration for the full 300 steps.
    On this data TENO5 amplifies last-bit differences (its stencil choice
    rests on smoothness indicators that are rounding noise in flat regions),
    so the bar is the reference's own reproducibility on the same problem:
    how far its result moves when only rounding changes -- its sources
    built with FMA contraction, and its run from the initial state with the
    last bit of every density flipped -- measured here on the CPU.  The
    device's distance from the oracle (itself bit-identical with the
    reference's build without contraction) is held within ten times the
    larger of the two, in relative L1 and Linf, plus 1e-11."""
    from oracle import bind as obd
    probs = _short_problems("teno5", 4)
    print(f"batch teno5 x{len(probs)}, 300 steps", flush=True)
    outs, _ = _run(probs, 300)
    for (kw, q), out in zip(probs, outs):
        n = kw["n"][0]
        desc = obd.make_desc(n=(n, 1, 1), active=(1, 0, 0), d=(kw["d"][0], 1.0, 1.0),
                             scheme="teno5", flux=kw["flux"], time="ssprk3")
        base, _ = obd.ref_run(desc, q, 300, 0.5)
        fma, _ = obd.ref_run(desc, q, 300, 0.5, fma=True)
        qb = np.asarray(q, dtype=np.float64).reshape(5, -1).copy()
        qb[0] = (qb[0].view(np.uint64) ^ np.uint64(1)).view(np.float64)
        ulp, _ = obd.ref_run(desc, qb, 300, 0.5)
        spread = [rel_errors(fma, base), rel_errors(ulp, base)]
        s1 = max(e[0].max() for e in spread)
        sinf = max(e[1].max() for e in spread)
        ref, _ = _oracle(kw, q, 300)
        d1, dinf = (e.max() for e in rel_errors(out, ref))
        print(f"teno5/{kw['flux']} n={n}: device vs oracle L1 {d1:.2e} Linf {dinf:.2e}; "
              f"reference spread L1 {s1:.2e} Linf {sinf:.2e}", flush=True)
        assert d1 <= 10.0 * s1 + 1e-11 and dinf <= 10.0 * sinf + 1e-11


def test_alone_equals_mixed_bitwise():
    probs = _mixed_problems()
    steps = 20
    mixed, mclocks = _run(probs, steps)
    for k in (0, 5, len(probs) - 3, len(probs) - 1):
        alone, aclocks = _run([probs[k]], steps)
        assert np.array_equal(alone[0], mixed[k]), f"problem {k} differs alone vs mixed"
        assert aclocks[0] == mclocks[k]


@pytest.mark.parametrize("graph,fused,tiles,occupancy", [
    (False, True, (127, 32), 8), (True, False, (127, 32), 8), (True, True, (63, 8), 2),
    (False, False, (255, 64), 1), (True, True, (255, 16), 32)])
def test_schedule_and_shape_do_not_change_results(graph, fused, tiles, occupancy):
    probs = _mixed_problems(count=6)
    base, bclocks = _run(probs, 12)
    other, oclocks = _run(probs, 12, graph=graph, fused=fused, tiles=tiles, occupancy=occupancy)
    for k in range(len(probs)):
        assert np.array_equal(base[k], other[k]), k
        assert bclocks[k] == oclocks[k]


def test_batch_matches_grid_path():
    kw, q = wl.riemann_1d(500, (1.0, 0.75, 1.0), (0.125, 0.0, 0.1), 0.4)
    kw.update(scheme="weno5", flux="hllc", cfl=0.5)
    outs, clocks = _run([(kw, q)], 30)
    with Grid(make_desc(n=(501, 1, 1), d=(1.0 / 500, 1.0, 1.0), scheme="weno5", flux="hllc",
                        time="ssprk3")) as g:
        g.upload(q)
        dts = g.run(30, 0.5)
        ref = g.download()
    assert_close(outs[0], ref, what="batch vs grid")
    assert abs(clocks[0][0] - dts.sum()) <= 1e-13 * dts.sum()


def test_step_limits_freeze_problems():
    probs = _mixed_problems(count=4)
    cap = sum(item_nodes(make_item(**kw)) for kw, _ in probs) + 64
    with Batch(cap, 16) as b:
        slots = [b.add(make_item(**kw), q, max_steps=5 + 3 * k) for k, (kw, q) in enumerate(probs)]
        b.step(40)
        steps = b.steps()
        for k, s in enumerate(slots):
            assert steps[s] == 5 + 3 * k
        outs = [b.download(s) for s in slots]
    for k, ((kw, q), out) in enumerate(zip(probs, outs)):
        ref, _ = _oracle(kw, q, 5 + 3 * k)
        assert_close(out, ref, what=f"frozen problem {k}")


def test_admission_by_footprint_and_reuse():
    kw1, q1 = wl.riemann_1d(300, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.5)
    kw1.update(scheme="weno5", flux="rusanov", cfl=0.5)
    kw2, q2 = wl.quadrant_2d(40, QUADS)
    kw2.update(scheme="weno5", flux="hllc", cfl=0.5)
    n1, n2 = item_nodes(make_item(**kw1)), item_nodes(make_item(**kw2))
    assert n1 == 307 and n2 == 46 * 46
    with Batch(2 * n1 + n2, 8) as b:
        a = b.add(make_item(**kw1), q1)
        c = b.add(make_item(**kw2), q2)
        a2 = b.add(make_item(**kw1), q1)
        assert b.try_add(make_item(**kw1), q1) is None   # arena full
        used, free, items = b.memory()
        assert (used, free, items) == (2 * n1 + n2, 0, 3)
        b.step(10)
        b.remove(a)
        b.remove(a2)
        assert b.memory()[1] == 2 * n1
        # a freed range is reused and the newcomer starts from step 0
        a3 = b.add(make_item(**kw1), q1)
        b.step(10)
        t3, k3 = b.clock(a3)
        tc, kc = b.clock(c)
        assert (k3, kc) == (10, 20)
        out1 = b.download(a3)
        out2 = b.download(c)
    ref1, _ = _oracle(kw1, q1, 10)
    ref2, _ = _oracle(kw2, q2, 20)
    assert_close(out1, ref1, what="readmitted 1d")
    assert_close(out2, ref2, what="resident 2d")


def test_autotune_leaves_problems_unchanged():
    probs = _mixed_problems(count=6)
    base, bclocks = _run(probs, 8)
    cap = sum(item_nodes(make_item(**kw)) for kw, _ in probs) + 64
    with Batch(cap, len(probs)) as b:
        slots = [b.add(make_item(**kw), q) for kw, q in probs]
        b.step(3)
        tuning = b.autotune(2)
        print(f"autotune: {tuning}", flush=True)
        assert tuning["candidates"] == 90 and tuning["ms_per_step"] > 0
        b.step(5)
        for k, s in enumerate(slots):
            assert np.array_equal(b.download(s), base[k]), k
            assert b.clock(s) == bclocks[k]


def test_faulted_problem_is_isolated():
    """A problem whose state is non-physical raises its own flags and is
    frozen; the others advance exactly as they would alone."""
    probs = _mixed_problems(count=4)
    base, bclocks = _run(probs, 6)
    kwb, qb = wl.riemann_1d(100, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.5)
    qb = qb.copy()
    qb[4, 40] = -1.0  # negative total energy: no valid sound speed
    kwb.update(scheme="weno5", flux="rusanov", cfl=0.5)
    cap = sum(item_nodes(make_item(**kw)) for kw, _ in probs) + item_nodes(make_item(**kwb)) + 64
    with Batch(cap, len(probs) + 1) as b:
        slots = [b.add(make_item(**kw), q) for kw, q in probs]
        bad = b.add(make_item(**kwb), qb)
        with pytest.raises(SonicError) as e:
            b.step(6)
        assert e.value.code == -3
        faults = b.faults()
        assert faults[bad] != 0
        assert all(faults[s] == 0 for s in slots)
        for k, s in enumerate(slots):
            assert np.array_equal(b.download(s), base[k]), k
            assert b.clock(s) == bclocks[k]


def test_batched_configuration_fails_where_the_reference_fails():
    """All 4096 problems of the batched configuration for its 300 steps: the
    problems the device finds non-physical are those on which the oracle
    (bit-identical to the reference's translation units) stops, as the
    reference's EOS closure, TENO5 cutoff and validateEquationState checks
    would stop it (the fixture tests/golden/batch_riemann1d_failures.json,
    written by tests/golden/make_batch_failures.py over every problem on the
    CPU); each of them is counted up to the step before its fault, as the
    oracle records."""
    import json
    from pathlib import Path
    ref = json.loads((Path(__file__).parent / "golden" /
                      "batch_riemann1d_failures.json").read_text())
    probs = [(kw, q) for _, kw, q in wl.batch_riemann_1d(ref["count"], seed=ref["seed"])]
    cap = sum(item_nodes(make_item(**kw)) for kw, _ in probs) + 64
    print(f"batched configuration: {len(probs)} problems, {ref['steps']} steps", flush=True)
    with Batch(cap, len(probs)) as b:
        b.set_tiles(255, 32)
        b.set_occupancy(16)
        b.set_schedule(False)
        slots = [b.add(make_item(**kw), q, max_steps=ref["steps"]) for kw, q in probs]
        with pytest.raises(SonicError) as e:
            b.step(ref["steps"])
        assert e.value.code == -3
        faults = b.faults()
        steps = b.steps()
    failed = sorted(k for k in range(len(probs)) if faults[slots[k]])
    print(f"device: {len(failed)} failed; oracle: {len(ref['failed'])}", flush=True)
    assert failed == ref["failed"]
    done = {r["index"]: r["steps_completed"] for r in ref["details"]}
    for k in range(len(probs)):
        expected = done.get(k, ref["steps"])
        assert steps[slots[k]] == expected, (k, steps[slots[k]], expected)


@pytest.mark.parametrize("fused", [False, True])
def test_mid_step_faults_leave_other_problems_bitwise_unchanged(fused):
    """Multi-tile problems of the batched configuration that turn
    non-physical part-way through a step (at a stage's validation or in a
    face flux's closure, after 2, 2 and 6 complete steps: the fixture's
    record of the oracle) under 256-lane x tiles, next to healthy 1D and 2D
    problems: each faulting problem stops at the oracle's step, and every
    healthy one advances bit for bit as it does without them -- the
    per-step liveness every stage kernel reads keeps the waves of a block in
    step when other blocks raise a fault within the same launch."""
    import json
    from pathlib import Path
    ref = json.loads((Path(__file__).parent / "golden" /
                      "batch_riemann1d_failures.json").read_text())
    done = {r["index"]: r["steps_completed"] for r in ref["details"]}
    picks = [2295, 3787, 2770]          # 6688, 4496, 2095 cells
    bad = {i: (kw, q) for i, kw, q in wl.batch_riemann_1d(ref["count"], seed=ref["seed"])
           if i in picks}
    probs = _mixed_problems(count=6)
    steps = 12
    base, bclocks = _run(probs, steps, fused=fused, tiles=(255, 32), occupancy=16)
    cap = (sum(item_nodes(make_item(**kw)) for kw, _ in probs) +
           sum(item_nodes(make_item(**kw)) for kw, _ in bad.values()) + 64)
    print(f"mid-step faults, fused={fused}", flush=True)
    with Batch(cap, len(probs) + len(bad)) as b:
        b.set_tiles(255, 32)
        b.set_occupancy(16)
        b.set_schedule(fused)
        slots = [b.add(make_item(**kw), q) for kw, q in probs]
        bslots = {i: b.add(make_item(**kw), q) for i, (kw, q) in bad.items()}
        with pytest.raises(SonicError) as e:
            b.step(steps)
        assert e.value.code == -3
        faults = b.faults()
        counts = b.steps()
        for i, s in bslots.items():
            assert faults[s] != 0 and counts[s] == done[i], (i, faults[s], counts[s])
        for k, s in enumerate(slots):
            assert faults[s] == 0
            assert np.array_equal(b.download(s), base[k]), k
            assert b.clock(s) == bclocks[k]
