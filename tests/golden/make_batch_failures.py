"""Which problems of the batched configuration the reference's algorithm
fails on: every problem of `batch_riemann_1d(4096)` (seed 4096, lengths
256..8192 cells, WENO5 / TENO5 x Rusanov / HLLC, SSPRK3, CFL 0.5) run for
300 steps by the oracle (oracle/liboracle.so, the C restatement, which the
tests hold bit-identical to the reference's own translation units).  A
problem fails where the oracle stops on a non-physical state, as the
reference's validateEquationState / TENO5 cutoff checks would.  Writes
batch_riemann1d_failures.json: the failing indices, each with the oracle's
flags and the number of steps it completed.

    python tests/golden/make_batch_failures.py [workers]

About an hour of CPU on 8 cores."""
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

COUNT, SEED, STEPS, CFL = 4096, 4096, 300, 0.5
OUT = Path(__file__).resolve().parent / "batch_riemann1d_failures.json"


def run_one(k):
    for i, kw, q in wl.batch_riemann_1d(COUNT, seed=SEED):
        if i == k:
            break
    n = kw["n"][0]
    desc = ob.make_desc(n=(n, 1, 1), active=(1, 0, 0), d=(kw["d"][0], 1.0, 1.0),
                        scheme=kw["scheme"], flux=kw["flux"], time="ssprk3")
    lib = ob.oracle()
    g = lib.sfo_grid_create(ctypes.byref(desc))
    try:
        a = np.ascontiguousarray(q, dtype=np.float64).reshape(5, -1).copy()
        lib.sfo_set_state(g, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        dts = np.zeros(STEPS)
        rc = lib.sfo_run(g, STEPS, CFL, 1e300, dts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc == 0:
            return None
        return {"index": k, "n": n, "scheme": kw["scheme"], "flux": kw["flux"],
                "flags": int(lib.sfo_error(g)), "steps_completed": int(np.count_nonzero(dts)) - 1}
    finally:
        lib.sfo_grid_destroy(g)


def main():
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    # longest problems first, so the pool finishes evenly
    order = sorted(range(COUNT), key=lambda k: -lengths[k])
    with Pool(workers) as pool:
        res = [r for r in pool.imap_unordered(run_one, order, chunksize=4) if r]
    res.sort(key=lambda r: r["index"])
    OUT.write_text(json.dumps({
        "workload": "batch_riemann_1d(4096, seed=4096): 1D Riemann problems, 256..8192 cells, "
                    "WENO5/TENO5 x Rusanov/HLLC, SSPRK3, CFL 0.5, 300 steps",
        "count": COUNT, "seed": SEED, "steps": STEPS,
        "checker": "oracle/liboracle.so (C restatement, bit-identical to the reference's TUs)",
        "failed": [r["index"] for r in res],
        "details": res}, indent=1) + "\n")
    print(f"{len(res)} of {COUNT} problems fail on the oracle -> {OUT}")


lengths = {i: kw["n"][0] for i, kw, _ in wl.batch_riemann_1d(COUNT, seed=SEED)}

if __name__ == "__main__":
    main()
