"""Which requests of the served configuration the reference's algorithm
fails on: the four-quadrant 2D Riemann requests of
`quadrant_requests(480, seed=5005)` (N log-uniform in [64, 1024],
50-500 steps, WENO5 / HLLC / SSPRK3, CFL 0.5, transmissive ends) run for
their own step counts by the oracle (oracle/liboracle.so, the C
restatement).  A request fails where the oracle stops on a non-physical
state, as the reference's flux closure and validateEquationState checks
would.

Checked: every request with N <= 256 (247 of the 480: both ways, so
that a request the device would fail and the oracle would not shows up),
plus the indices given on the command line (those the device flagged,
whatever their size).  Writes served_quadrant_failures.json: the checked
indices, the failing ones, each with the oracle's flags and the steps it
completed.

    python tests/golden/make_served_failures.py [workers] [index ...]

About 20 minutes of CPU on 6 cores."""
import ctypes
import json
import sys
from multiprocessing import Pool
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import bind as ob  # noqa: E402
from sonicsolver_amd import workloads as wl  # noqa: E402

COUNT, SEED, SMALL = 480, 5005, 256
OUT = Path(__file__).resolve().parent / "served_quadrant_failures.json"


def _request(k):
    for i, kw, steps, make in wl.quadrant_requests(COUNT, seed=SEED):
        if i == k:
            return kw, steps, make()
    raise IndexError(k)


def run_one(k):
    kw, steps, q = _request(k)
    n = kw["n"]
    desc = ob.make_desc(n=(n[0], n[1], 1), active=(1, 1, 0), d=(kw["d"][0], kw["d"][1], 1.0),
                        scheme=kw["scheme"], flux=kw["flux"], time="ssprk3")
    lib = ob.oracle()
    g = lib.sfo_grid_create(ctypes.byref(desc))
    try:
        a = np.ascontiguousarray(q, dtype=np.float64).reshape(5, -1).copy()
        lib.sfo_set_state(g, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        dts = np.zeros(steps)
        rc = lib.sfo_run(g, steps, kw["cfl"], 1e300,
                         dts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc == 0:
            return k, None
        return k, {"index": k, "n": n[0], "steps": steps, "flags": int(lib.sfo_error(g)),
                   "steps_completed": int(np.count_nonzero(dts)) - 1}
    finally:
        lib.sfo_grid_destroy(g)


def main():
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 6
    extra = sorted({int(a) for a in sys.argv[2:]})
    sizes = {i: (kw["n"][0], steps) for i, kw, steps, _ in wl.quadrant_requests(COUNT, seed=SEED)}
    checked = sorted({i for i, (n, _) in sizes.items() if n <= SMALL} | set(extra))
    # costliest first, so the pool finishes evenly
    order = sorted(checked, key=lambda k: -sizes[k][0] ** 2 * sizes[k][1])
    with Pool(workers) as pool:
        res = [r for _, r in pool.imap_unordered(run_one, order) if r]
    res.sort(key=lambda r: r["index"])
    OUT.write_text(json.dumps({
        "workload": f"quadrant_requests({COUNT}, seed={SEED}): four-quadrant 2D Riemann "
                    "requests, N log-uniform [64, 1024], 50-500 steps, WENO5/HLLC/SSPRK3, "
                    "CFL 0.5",
        "count": COUNT, "seed": SEED,
        "checker": "oracle/liboracle.so (C restatement, bit-identical to the reference's TUs)",
        "checked": checked, "checked_rule": f"every request with N <= {SMALL}, plus {extra}",
        "failed": [r["index"] for r in res],
        "details": res}, indent=1) + "\n")
    print(f"{len(res)} of {len(checked)} checked requests fail on the oracle -> {OUT}")


if __name__ == "__main__":
    main()
