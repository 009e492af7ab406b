"""Served solves: requests submitted while others run are admitted as arena
memory and slots free up; every result equals the same problem solved alone
(bit for bit) and matches the oracle (relative L1 / Linf <= 1e-11)."""
import threading
import time

import numpy as np
import pytest

from oracle import bind as ob
from sonicsolver_amd import SonicError, workloads as wl
from sonicsolver_amd.batch import Batch, item_nodes, make_item
from sonicsolver_amd.serve import Server
from tests.parity import assert_close

pytestmark = pytest.mark.gpu


def _requests():
    reqs = []
    for _, kw, steps, make in wl.quadrant_requests(10, seed=77, lo=16, hi=96, steps=(8, 40)):
        reqs.append((kw, steps, make()))
    for i, kw, q in wl.batch_riemann_1d(6, seed=9, lo=64, hi=600):
        reqs.append((kw, 10 + 3 * i, q))
    return reqs


def _alone(kw, steps, q):
    with Batch(item_nodes(make_item(**kw)) + 64, 1) as b:
        s = b.add(make_item(**kw), q)
        b.step(steps)
        return b.download(s), b.clock(s)


def test_served_equals_alone_and_oracle():
    reqs = _requests()
    biggest = max(item_nodes(make_item(**kw)) for kw, _, _ in reqs)
    # room for about three problems: admission must wait for slots to free up
    budget = 160 * 3 * biggest
    results = {}
    with Server(budget, max_items=6, steps_per_tick=7) as srv:
        tickets = []

        def client(part):
            for k in part:
                kw, steps, q = reqs[k]
                tickets.append((k, srv.submit(make_item(**kw), q, steps)))
                time.sleep(0.002)

        threads = [threading.Thread(target=client, args=(range(j, len(reqs), 2),))
                   for j in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        srv.drain(120)
        st = srv.stats()
        for k, t in tickets:
            results[k] = srv.result(t)
    print(f"served {st['completed']} in {st['ticks']} ticks, peak {st['peak_items']} resident, "
          f"p50 {st['p50_ms']:.1f} ms p99 {st['p99_ms']:.1f} ms", flush=True)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_nodes"] <= st["arena_nodes"]
    for k, (kw, steps, q) in enumerate(reqs):
        out, t, nsteps, lat = results[k]
        assert nsteps == steps
        alone, (ta, ka) = _alone(kw, steps, q)
        assert np.array_equal(out, alone), f"request {k} differs from the alone solve"
        assert t == ta and ka == steps
        if k % 3 == 0:
            n = kw["n"]
            two_d = n[1] > 1
            desc = ob.make_desc(n=(n[0], n[1] if two_d else 1, 1),
                                active=(1, 1 if two_d else 0, 0),
                                d=(kw["d"][0], kw["d"][1] if two_d else 1.0, 1.0),
                                scheme=kw["scheme"], flux=kw["flux"], time="ssprk3")
            ref, _ = ob.run(desc, q, steps, kw["cfl"])
            assert_close(out, ref, what=f"served request {k} n={n}")


def test_oversized_request_rejected():
    kw, q = wl.quadrant_2d(64, [(1, 0, 0, 1)] * 4)
    kw.update(scheme="weno5", flux="hllc", cfl=0.5)
    with Server(160 * 1000, max_items=2) as srv:
        with pytest.raises(SonicError) as e:
            srv.submit(make_item(**kw), q, 10)
        assert e.value.code == -4


def test_served_configuration_fails_where_the_reference_fails():
    """The served configuration's requests (quadrant_requests(480, seed=5005),
    each for its own step count) that the oracle checked on the CPU (every
    request with N <= 256, plus every request the device flagged in the
    bench's served run: the fixture tests/golden/served_quadrant_failures.json,
    written by tests/golden/make_served_failures.py): the device finds
    exactly the oracle's failing requests non-physical, each after the
    oracle's number of completed steps.  Run through one batch, which is
    what the request loop advances (served results equal batch results bit
    for bit, test_served_equals_alone_and_oracle)."""
    import json
    from pathlib import Path
    ref = json.loads((Path(__file__).parent / "golden" /
                      "served_quadrant_failures.json").read_text())
    checked = set(ref["checked"])
    reqs = [(i, kw, steps, make) for i, kw, steps, make in
            wl.quadrant_requests(ref["count"], seed=ref["seed"]) if i in checked]
    cap = sum(item_nodes(make_item(**kw)) for _, kw, _, _ in reqs) + 64
    longest = max(steps for _, _, steps, _ in reqs)
    print(f"served configuration: {len(reqs)} checked requests, up to {longest} steps", flush=True)
    with Batch(cap, len(reqs)) as b:
        slots = {i: b.add(make_item(**kw), make(), max_steps=steps) for i, kw, steps, make in reqs}
        with pytest.raises(SonicError) as e:
            b.step(longest)
        assert e.value.code == -3
        faults = b.faults()
        done = b.steps()
    failed = sorted(i for i in slots if faults[slots[i]])
    print(f"device: {len(failed)} failed; oracle: {len(ref['failed'])}", flush=True)
    assert failed == ref["failed"]
    completed = {r["index"]: r["steps_completed"] for r in ref["details"]}
    for i, kw, steps, _ in reqs:
        assert done[slots[i]] == completed.get(i, steps), (i, done[slots[i]], completed.get(i, steps))
