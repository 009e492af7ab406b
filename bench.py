#!/usr/bin/env python3
"""Benchmark of the explicit Euler hot path on one MI355X.

Workload: 3D Euler on the periodic cube [0, 2 pi)^3 with 512^3 nodes, a
seeded Taylor-Green-type initial state, characteristic WENO5-JS
reconstruction, HLLC flux, SSPRK3, CFL 0.4 (the configuration the headline
metric is quoted on).  A step is one physical SSPRK3 step over the whole
grid; the state stays resident in HBM for the whole run.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  `value` is Mcell-updates/s (one cell-update = one
node advanced by one full SSPRK3 step).  `roofline` is the dominant kernel's
algorithmic FP64 operation rate, timed live with HIP events on the grid's
stream, against MI355X's FP64 peak (78.6 TFLOP/s); `hbm` is the step's
compulsory traffic rate against 8 TB/s; `cpu_baseline` times the reference's
own translation units (oracle/_ref/libsfref.so) on a bounded sample of the
same workload, serially and on every host core as z slabs.
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mcell-updates/s FP64 WENO5+RK3 (% HBM peak); batched solves/s + p99 latency"
HBM_PEAK_GBPS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# Host threads a one-GPU job may use on the machines this runs on (its CPU
# share of a shared multi-GPU host, whatever the host's own core count is).
HOST_THREAD_SHARE = 16


def cpu_baseline(n: int, steps: int):
    """The reference's own code (harness over its translation units) on an
    n^3 sample of the same workload: serially, and on every host core the
    job may use (the one-GPU share of HOST_THREAD_SHARE threads) as z
    slabs, one thread per slab, with the halo planes exchanged before each
    ghost fill and the dt reduced over slabs (the reference's
    parallel.split decomposition, in shared memory; bit-identical to the
    serial run).  Both time the steps alone.  Falls back to the C
    restatement, serial, where the harness was not built."""
    from oracle import bind as ob
    from sonicsolver_amd import workloads as wl

    kw, q = wl.tgv_3d(n)
    desc = ob.make_desc(**kw)
    if not ob.ref_available():
        t0 = time.perf_counter()
        ob.run(desc, q, steps, 0.4)
        dt = time.perf_counter() - t0
        return {"value": round(n ** 3 * steps / dt / 1e6, 4), "unit": "Mcell-updates/s",
                "cores": 1, "kind": "port",
                "sample": f"3D periodic TGV {n}^3, WENO5-JS/HLLC/SSPRK3 CFL 0.4, {steps} steps, "
                          f"C restatement oracle/liboracle.so, serial, {dt:.1f} s"}
    _, _, dt = ob.ref_run_slabs(desc, q, steps, 0.4, 1)
    out = {
        "value": round(n ** 3 * steps / dt / 1e6, 4),
        "unit": "Mcell-updates/s",
        "cores": 1,
        "kind": "reference",
        "sample": f"3D periodic TGV {n}^3, WENO5-JS/HLLC/SSPRK3 CFL 0.4, {steps} steps, "
                  f"reference translation units via oracle/_ref/libsfref.so, serial, "
                  f"{dt:.1f} s of steps",
    }
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    halo = 3
    cores = max(1, min(HOST_THREAD_SHARE, avail, n // halo))
    psteps = 6 * steps
    _, _, pdt = ob.ref_run_slabs(desc, q, psteps, 0.4, cores)
    out["all_cores"] = {
        "value": round(n ** 3 * psteps / pdt / 1e6, 4),
        "unit": "Mcell-updates/s",
        "cores": cores,
        "kind": "reference",
        "mode": f"{cores} z slabs, one thread each, halo planes exchanged before every ghost "
                f"fill, dt reduced over slabs (mpirun is not on the image)",
        "host_cores_visible": os.cpu_count(),
        "host_cores_usable": avail,
        "why_not_more": f"a one-GPU job may use {HOST_THREAD_SHARE} host threads on this "
                        f"machine, whatever the host's core count",
        "sample": f"the same {n}^3 field, {psteps} steps, {pdt:.1f} s of steps",
    }
    return out


def _plain_case(kw, q, steps):
    """One problem on its own grid through the single-case path (one launch
    sequence per step); returns (state, seconds spent in the run)."""
    from sonicsolver_amd import Grid, make_desc
    n = kw["n"]
    two_d = len(n) > 1 and n[1] > 1
    desc = make_desc(n=(n[0], n[1] if two_d else 1, 1), active=(1, 1 if two_d else 0, 0),
                     d=(kw["d"][0], kw["d"][1] if two_d else 1.0, 1.0), scheme=kw["scheme"],
                     flux=kw["flux"], time="ssprk3")
    with Grid(desc) as g:
        g.upload(q)
        g.synchronize()
        t0 = time.perf_counter()
        g.run(steps, kw["cfl"], record_dt=False)
        g.synchronize()
        el = time.perf_counter() - t0
        return g.download(), el


def _rel_err(a, b):
    a = np.asarray(a).reshape(5, -1)
    b = np.asarray(b).reshape(5, -1)
    d = np.abs(a - b)
    return float(max((d[v].sum() / max(np.abs(b[v]).sum(), 1e-300)) for v in range(5))), \
        float(max((d[v].max() / max(np.abs(b[v]).max(), 1e-300)) for v in range(5)))


def batched_run(count: int, steps: int, sample: int, launch: int = 1, shape=None):
    """Independent 1D Riemann problems of uneven length (256..8192 cells,
    mixed WENO5/TENO5 and Rusanov/HLLC, CFL 0.5 each) advanced together by
    the persistent batch kernels for `steps` steps each, against the same GPU
    solving them one case at a time."""
    from sonicsolver_amd import SonicError, flops, workloads as wl
    from sonicsolver_amd.batch import Batch, item_nodes, make_item
    probs = [(kw, q) for _, kw, q in wl.batch_riemann_1d(count)]
    cells = sum(kw["n"][0] - 1 for kw, _ in probs)
    cap = sum(item_nodes(make_item(**kw)) for kw, _ in probs) + 64
    picks = list(range(0, count, max(1, count // sample)))[:sample]
    with Batch(cap, count) as b:
        slots = [b.add(make_item(**kw), q, max_steps=steps) for kw, q in probs]
        if shape:
            b.set_tiles(shape[0], shape[1])
            b.set_occupancy(shape[2])
            b.set_schedule(bool(shape[3]))
            tuning = {"xlen": shape[0], "ylen": shape[1], "blocks_per_cu": shape[2],
                      "fused": shape[3], "given": True}
            log(f"batched: {count} problems admitted ({cells} cells), shape {tuning}")
        else:
            log(f"batched: {count} problems admitted ({cells} cells), autotuning")
            tuning = b.autotune(3)
            log(f"batched: tuned {tuning}; timing {steps} steps")
        t0 = time.perf_counter()
        try:
            b.step(steps, graph=launch)
        except SonicError as e:
            if e.code != -3:
                raise
        elapsed = time.perf_counter() - t0
        # a problem that reaches a non-physical state is frozen with its
        # flags while the others run on (the reference's CPU path stops on
        # the same problems)
        faults = b.faults()[:count]
        failed = [k for k in range(count) if faults[slots[k]]]
        done = b.steps()[:count]
        assert all(done[slots[k]] == steps for k in range(count) if not faults[slots[k]])
        picks = [k for k in picks if not faults[slots[k]]]
        batch_out = {k: b.download(slots[k]) for k in picks}
    log(f"batched: {elapsed:.3f} s; plain single-case path on {len(picks)} of the problems")
    plain_s = 0.0
    plain_cells = 0
    worst = (0.0, 0.0)
    for k in picks:
        kw, q = probs[k]
        out, el = _plain_case(kw, q, steps)
        plain_s += el
        plain_cells += kw["n"][0] - 1
        e = _rel_err(batch_out[k], out)
        worst = (max(worst[0], e[0]), max(worst[1], e[1]))
    solved = count - len(failed)
    solved_cells = cells - sum(probs[k][0]["n"][0] - 1 for k in failed)
    # algorithmic FP64 work of the solved problems (each recipe's count)
    fail_set = set(failed)
    flop = sum((kw["n"][0] - 1) * steps * flops.flops_per_cell_step(1, kw["scheme"], kw["flux"], "ssprk3")
               for k, (kw, _) in enumerate(probs) if k not in fail_set)
    solves = solved / elapsed
    plain_solves = len(picks) / plain_s
    return {
        "workload": f"{count} independent 1D Riemann problems, 256-8192 cells (log-uniform), "
                    f"WENO5/TENO5 x Rusanov/HLLC, SSPRK3, per-problem CFL 0.5 dt, {steps} steps "
                    "each, FP64",
        "solves_per_s": round(solves, 2),
        "mcell_updates_per_s": round(solved_cells * steps / elapsed / 1e6, 2),
        "fp64_frac": round(flop / elapsed / (FP64_PEAK_TFLOPS * 1e12), 4),
        "seconds": round(elapsed, 4),
        "cells": cells,
        "solved": solved,
        "failed_nonphysical": len(failed),
        "failed_problems": failed,
        "oracle_check": _failure_agreement(failed, count, steps),
        "launch": {0: "direct", 1: "hipGraph replay", 2: "automatic"}[launch],
        "tuning": tuning,
        "plain_single_case": {
            "solves_per_s": round(plain_solves, 2),
            "mcell_updates_per_s": round(plain_cells * steps / plain_s / 1e6, 2),
            "sample": f"{len(picks)} of the problems, each on its own grid, sf_run per case",
        },
        "speedup_vs_plain": round(solves / plain_solves, 2),
        "max_rel_diff_vs_plain": {"l1": worst[0], "linf": worst[1]},
    }


def _failure_agreement(failed, count, steps):
    """The problems the device found non-physical against the ones the
    reference's algorithm fails on (the oracle run over every problem of the
    batched configuration on the CPU, tests/golden/make_batch_failures.py,
    stored as a fixture): both counts and whether the sets agree."""
    path = ROOT / "tests" / "golden" / "batch_riemann1d_failures.json"
    if not path.exists():
        return None
    ref = json.loads(path.read_text())
    if ref.get("count") != count or ref.get("steps") != steps:
        return None
    oracle_failed = sorted(int(k) for k in ref["failed"])
    return {"device_failed": len(failed), "oracle_failed": len(oracle_failed),
            "same_problems": sorted(failed) == oracle_failed,
            "fixture": str(path.relative_to(ROOT))}


def _served_failure_agreement(failed, count, seed):
    """The served requests the device found non-physical against the
    reference's algorithm on the CPU (tests/golden/make_served_failures.py:
    the oracle over every request with N <= 256 and over each request the
    device flagged, stored as a fixture): within the checked requests both
    sets must agree, and every flagged request must fail on the oracle."""
    path = ROOT / "tests" / "golden" / "served_quadrant_failures.json"
    if not path.exists():
        return None
    ref = json.loads(path.read_text())
    if ref.get("count") != count or ref.get("seed") != seed:
        return None
    checked = set(int(k) for k in ref["checked"])
    oracle_failed = sorted(int(k) for k in ref["failed"])
    device_checked = sorted(k for k in failed if k in checked)
    return {"device_failed": len(failed), "oracle_failed": len(oracle_failed),
            "checked_requests": len(checked),
            "same_requests": device_checked == oracle_failed and len(device_checked) == len(failed),
            "fixture": str(path.relative_to(ROOT))}


def configs_1d_2d():
    """The two single-grid configurations besides the 3D headline: the 1000-
    cell shock tube (WENO5/Rusanov/SSPRK3, CFL 0.5, 200 steps) and the 2048^2
    periodic isentropic vortex with density noise (TENO5/HLLC/SSPRK3, CFL
    0.5, 100 steps), timed with the state resident on the device; the vortex
    also reports its error against the exact translated vortex."""
    from sonicsolver_amd import Grid, make_desc, workloads as wl
    out = {}
    kw, q = wl.sod_1d(1000)
    with Grid(make_desc(**kw)) as g:
        g.upload(q)
        g.run(10, 0.5)
        g.upload(q)
        g.synchronize()
        t0 = time.perf_counter()
        g.run(200, 0.5)
        g.synchronize()
        el = time.perf_counter() - t0
    out["sod_1d_1000"] = {"steps": 200, "seconds": round(el, 5),
                          "ms_per_step": round(el / 200 * 1e3, 4),
                          "mcell_updates_per_s": round(1001 * 200 / el / 1e6, 3)}
    n = 2048
    kw, q = wl.vortex_2d(n)
    with Grid(make_desc(**kw)) as g:
        g.upload(q)
        g.run(2, 0.5)
        g.upload(q)
        g.synchronize()
        t0 = time.perf_counter()
        dts = g.run(100, 0.5)
        g.synchronize()
        el = time.perf_counter() - t0
        res = g.download()
    t = float(np.sum(dts))
    h = 10.0 / n
    y, x = np.meshgrid(np.arange(n) * h, np.arange(n) * h, indexing="ij")
    rho, u, v, p = wl.vortex_state(x, y, t)
    err = np.abs(res[0] - rho.ravel())
    out["vortex_2d_2048"] = {"steps": 100, "seconds": round(el, 4),
                             "ms_per_step": round(el / 100 * 1e3, 4),
                             "mcell_updates_per_s": round(n * n * 100 / el / 1e6, 2),
                             "time": t,
                             "density_error_vs_exact": {"l1": float(err.mean()),
                                                        "linf": float(err.max())},
                             "note": "the initial density carries 1e-3 seeded noise, so the "
                                     "error against the noise-free exact vortex is O(1e-3)"}
    return out


def recipes_run(n: int, steps: int):
    """Every registered convection recipe (WENO3 / WENO5 / TENO5 / WENO7 x
    Steger-Warming / Rusanov / Lax-Friedrichs / Roe / Lax-Wendroff, plus
    HLLC) on the same
    n^3 periodic Taylor-Green-type field with SSPRK3 at CFL 0.4: steps timed
    after one warm-up step, the state resident on the device.  Reported as
    cell updates per second and as FP64 rate against the per-recipe
    operation count."""
    from sonicsolver_amd import Grid, make_desc, workloads as wl, flops
    kw, q = wl.tgv_3d(n)
    out = {}
    for scheme in ("weno3", "weno5", "teno5", "weno7"):
        for flux in ("steger", "rusanov", "laxfriedrichs", "roe", "laxwendroff", "hllc"):
            kw2 = dict(kw, scheme=scheme, flux=flux)
            with Grid(make_desc(**kw2)) as g:
                g.upload(q)
                g.run(1, 0.4)
                g.synchronize()
                t0 = time.perf_counter()
                g.run(steps, 0.4)
                g.synchronize()
                el = time.perf_counter() - t0
                fl = g.flags()
            per = flops.flops_per_cell_step(3, scheme, flux, "ssprk3")
            tf = per * n ** 3 * steps / el / 1e12
            out[f"{scheme}_{flux}"] = {"mcell_updates_per_s": round(n ** 3 * steps / el / 1e6, 1),
                                       "ms_per_step": round(el / steps * 1e3, 3),
                                       "flops_per_cell_step": per,
                                       "achieved_tflops": round(tf, 2),
                                       "fp64_frac": round(tf / FP64_PEAK_TFLOPS, 4),
                                       "device_flags": fl}
            log(f"recipe {scheme}/{flux}: {out[f'{scheme}_{flux}']['mcell_updates_per_s']} Mcell/s")
    return {"workload": f"3D Euler {n}^3 periodic Taylor-Green-type field, SSPRK3, CFL 0.4, "
                        f"FP64, {steps} timed steps per recipe", "recipes": out}


def _serve_pass(reqs, rate, budget_gb, seed, keep, launch=2, shape=None):
    """One pass of the request stream through a fresh server: Poisson
    arrivals at `rate`/s submitted from a client thread; returns the
    server's statistics, the wall time and the results of the `keep`
    requests."""
    import threading
    from sonicsolver_amd import SonicError, workloads as wl
    from sonicsolver_amd.batch import make_item
    from sonicsolver_amd.serve import Server
    count = len(reqs)
    gaps = -np.log(1.0 - wl.uniform01(seed, count)) / rate
    arrivals = np.cumsum(gaps) - gaps[0]
    tickets = [None] * count
    with Server(int(budget_gb * 2**30), max_items=1024, steps_per_tick=25) as srv:
        srv.set_launch(launch)
        if shape:
            srv.set_shape(*shape)
        # warm the kernels and the graph path on one small request
        kw, st, q = min(reqs, key=lambda r: r[0]["n"][0])
        warm = srv.submit(make_item(**kw), q, 2)
        srv.wait(warm, 60)
        srv.result(warm)
        srv.reset_stats()

        def submitter():
            t0 = time.perf_counter()
            for k, (kw, st, q) in enumerate(reqs):
                delay = arrivals[k] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                tickets[k] = srv.submit(make_item(**kw), q, st)

        th = threading.Thread(target=submitter)
        t0 = time.perf_counter()
        th.start()
        th.join()
        srv.drain(600)
        wall = time.perf_counter() - t0
        stats = srv.stats()
        outs, failed = {}, []
        for k, t in enumerate(tickets):
            try:
                out, tt, nst, lat = srv.result(t)
            except SonicError as e:
                # a request whose state turns non-physical fails alone; the
                # reference's CPU path stops on the same problems
                if e.code != -3:
                    raise
                failed.append(k)
                continue
            if k in keep:
                outs[k] = out
    return stats, wall, outs, failed


def served_run(count: int, rate: float, budget_gb: float, sample: int, seed: int = 5005,
               launch: int = 2, shape=None):
    """Served stream of four-quadrant 2D Riemann problems (N log-uniform in
    [64, 1024], 50-500 steps, WENO5/HLLC/SSPRK3) arriving as a Poisson
    process; requests are admitted into the device arena by footprint as
    memory frees up (continuous batching).  A first pass offers `rate`
    requests/s (well above what the GPU sustains, so the queue stays full
    and the ramp-up and the final drain are a small part of the run) and
    measures solves/s from the first submission to the last result; a
    second pass offers 70% of that and measures the latency percentiles."""
    from sonicsolver_amd import flops, workloads as wl
    reqs = [(kw, steps, make()) for _, kw, steps, make in wl.quadrant_requests(count, seed=seed)]
    cells = sum(kw["n"][0] * kw["n"][1] * st for kw, st, _ in reqs)
    picks = list(range(0, count, max(1, count // sample)))[:sample]
    log(f"served: {count} requests at {rate}/s (Poisson), arena {budget_gb} GiB")
    sat, sat_wall, outs, failed = _serve_pass(reqs, rate, budget_gb, seed + 1, set(picks),
                                              launch, shape)
    moderate = round(0.7 * sat["solves_per_s"], 2)
    log(f"served: {sat['solves_per_s']:.1f} solves/s at saturation; latency pass at {moderate}/s")
    lat, lat_wall, _, _ = _serve_pass(reqs, moderate, budget_gb, seed + 2, set(), launch, shape)
    picks = [k for k in picks if k in outs]
    plain_s = 0.0
    worst = (0.0, 0.0)
    rates = []   # (N, cell-updates, seconds) of each plain solve
    for k in picks:
        kw, st, q = reqs[k]
        out, el = _plain_case(kw, q, st)
        plain_s += el
        rates.append((kw["n"][0], kw["n"][0] * kw["n"][1] * st, el))
        e = _rel_err(outs[k], out)
        worst = (max(worst[0], e[0]), max(worst[1], e[1]))
    plain_solves = len(picks) / plain_s
    big = sorted(rates)[-max(1, len(rates) // 4):]   # the largest quarter of the sample
    return {
        "workload": f"{count} four-quadrant 2D Riemann requests, N log-uniform [64,1024], "
                    "50-500 steps, WENO5/HLLC/SSPRK3 FP64, Poisson arrivals",
        "solves_per_s": round(sat["solves_per_s"], 2),
        "saturation_rate_per_s": rate,
        "mcell_updates_per_s": round(cells / sat_wall / 1e6, 2),
        "fp64_frac": round(cells * flops.flops_per_cell_step(2, "weno5", "hllc", "ssprk3")
                           / sat_wall / (FP64_PEAK_TFLOPS * 1e12), 4),
        "gpu_busy_frac": round(sat["busy_s"] / sat_wall, 3),
        "admit_frac": round(sat["admit_s"] / sat_wall, 3),
        "retire_frac": round(sat["retire_s"] / sat_wall, 3),
        "ticks": sat["ticks"],
        "peak_resident": sat["peak_items"],
        "latency_rate_per_s": moderate,
        "p50_ms": round(lat["p50_ms"], 2),
        "p99_ms": round(lat["p99_ms"], 2),
        "mean_ms": round(lat["mean_ms"], 2),
        "completed": sat["completed"],
        "failed_nonphysical": len(failed),
        "failed_requests": failed,
        "oracle_check": _served_failure_agreement(failed, count, seed),
        "plain_single_case": {
            "solves_per_s": round(plain_solves, 2),
            "mcell_updates_per_s": round(sum(r[1] for r in rates) / plain_s / 1e6, 2),
            "largest_quarter": {
                "n": [r[0] for r in big],
                "mcell_updates_per_s": round(sum(r[1] for r in big) / sum(r[2] for r in big) / 1e6, 2),
            },
            "sample": f"{len(picks)} of the requests, each on its own grid, back to back",
        },
        "speedup_vs_plain": round(sat["solves_per_s"] / plain_solves, 2),
        "max_rel_diff_vs_plain": {"l1": worst[0], "linf": worst[1]},
    }


def traffic_measurement(kernel: str):
    """HBM bytes per launch of `kernel` measured with rocprofv3 PMC counters
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size) and
    committed under profiles/; None where no measurement exists."""
    path = ROOT / "profiles" / "traffic_3d512.json"
    if not path.exists():
        return None, None
    data = json.loads(path.read_text())
    entry = data.get("kernels", {}).get(kernel)
    if not entry:
        return None, None
    return entry.get("bytes_per_launch"), str(path.relative_to(ROOT))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512, help="nodes per axis (default 512)")
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--cpu-sample", type=int, default=128, help="n of the n^3 CPU sample")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched", type=int, default=4096,
                    help="problems in the batched run (0 skips it)")
    ap.add_argument("--batched-steps", type=int, default=300)
    ap.add_argument("--batched-launch", choices=["graph", "direct"], default="graph",
                    help="replay one captured hipGraph per step, or launch each step's kernels")
    ap.add_argument("--batched-shape", default=None,
                    help="XLEN,YLEN,BLOCKS_PER_CU,FUSED instead of autotuning")
    ap.add_argument("--served-launch", choices=["auto", "direct"], default="auto",
                    help="the request loop's launches: graphs where they pay, or direct")
    ap.add_argument("--served-shape", default=None,
                    help="XLEN,YLEN,BLOCKS_PER_CU,FUSED of the request loop's batch "
                         "(default: the server's own)")
    ap.add_argument("--served", type=int, default=480,
                    help="requests in the served run (0 skips it)")
    ap.add_argument("--served-rate", type=float, default=120.0,
                    help="Poisson arrivals per s of the saturation pass (well above what "
                         "the GPU sustains, so the queue stays full)")
    ap.add_argument("--served-budget-gb", type=float, default=24.0)
    ap.add_argument("--plain-sample", type=int, default=32,
                    help="problems solved one case at a time for the plain comparison")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the 1D shock tube and 2D vortex configurations")
    ap.add_argument("--recipes", type=int, default=256,
                    help="n of the n^3 grid every convection recipe is timed on (0 skips)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device copy of known size (2 GiB read + 2 GiB written, "
                         "5 times) to calibrate HBM traffic counters, then exit")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = 1 << 28
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": n * 8, "bytes_written": n * 8,
                          "launches": 5}))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: sonicsolver_amd measures the explicit hot path on a single MI355X "
              "only (multi-GPU sharding is a design note, not a measured configuration). "
              "Re-run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from sonicsolver_amd import Grid, make_desc, workloads as wl, flops
    from sonicsolver_amd._native import KERNELS

    n = args.n
    scheme, flux, tscheme = "weno5", "hllc", "ssprk3"
    log(f"generating the {n}^3 initial state")
    kw, q = wl.tgv_3d(n, scheme=scheme, flux=flux, time=tscheme)
    cells = n ** 3
    with Grid(make_desc(**kw)) as g:
        log(f"grid allocated: {g.device_bytes() / 2**30:.1f} GiB of device memory")
        g.upload(q)
        mass0 = float(q[0].sum())
        del q
        log(f"warm-up: {args.warmup} steps")
        if args.warmup:
            g.run(args.warmup, args.cfl)
        g.synchronize()
        g.profile(True)
        log(f"timed: {args.steps} steps")
        t0 = time.perf_counter()
        dts = g.run(args.steps, args.cfl)
        g.synchronize()
        elapsed = time.perf_counter() - t0
        prof = {k: g.profile_read(v) for k, v in KERNELS.items()}
        g.profile(False)
        flags = g.flags()
        out = g.download()
        mass1 = float(out[0].sum())
        del out

    ms_per_step = elapsed / args.steps * 1e3
    value = cells * args.steps / elapsed / 1e6

    # dominant kernel: the sweep with the largest summed duration
    sweeps = {0: "sweep_x", 1: "sweep_y", 2: "sweep_z"}
    axis, name = max(sweeps.items(), key=lambda kv: prof[kv[1]][0])
    total_ms, launches = prof[name]
    avg_ms = total_ms / max(launches, 1)
    final = axis == 0  # the x sweep applies the stage update
    fl = flops.sweep_flops((n, n, n), axis, scheme, flux, final=final, dim=3, time=tscheme,
                           fold_dt=1.0 / 3.0 if final else 0.0)
    achieved_tf = fl / (avg_ms * 1e-3) / 1e12
    kernel_names = {"sweep_x": "the x sweep: k_sweep_xl on grids with many x lines (512^3), "
                               "else k_sweep_x (x faces + stage update + CFL fold)",
                    "sweep_y": "k_sweep_march<1> (y faces, accumulates S)",
                    "sweep_z": "k_sweep_march<2> (z faces, writes S)"}
    traffic, traffic_src = traffic_measurement(name)

    step_flops = flops.flops_per_cell_step(3, scheme, flux, tscheme) * cells
    hbm_gbps = flops.BYTES_PER_CELL_STEP[tscheme] * cells * args.steps / elapsed / 1e9

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Taylor-Green-type initial state, splitmix64 phases)",
        "config": {
            "workload": f"3D Euler {n}^3 periodic [0,2pi)^3 Taylor-Green-type field, "
                        "characteristic WENO5-JS + HLLC + SSPRK3, CFL 0.4, FP64",
            "grid": [n, n, n], "cells": cells, "scheme": scheme, "flux": flux,
            "time": tscheme, "cfl": args.cfl,
        },
        "roofline": {
            "bound": "fp64-valu",
            "pipe": "FP64 vector ALU: no dense contraction on this path, so neither the HBM "
                    "nor the matrix-core roof applies; the peak is MI355X's FP64 figure, "
                    "which is the same for vector and matrix FP64",
            "kernel": kernel_names[name],
            "achieved": round(achieved_tf, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tf / FP64_PEAK_TFLOPS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "flops_per_launch": fl,
            "avg_launch_ms": round(avg_ms, 4),
            "launches": launches,
        },
        "hbm": {
            "algorithmic_bytes_per_cell_step": flops.BYTES_PER_CELL_STEP[tscheme],
            "achieved": round(hbm_gbps, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(hbm_gbps / HBM_PEAK_GBPS, 4),
        },
        "step_flops": {
            "per_cell_step": flops.flops_per_cell_step(3, scheme, flux, tscheme),
            "achieved_tflops": round(step_flops * args.steps / elapsed / 1e12, 3),
            "frac": round(step_flops * args.steps / elapsed / 1e12 / FP64_PEAK_TFLOPS, 4),
        },
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in prof.items()},
        "checks": {"device_flags": flags, "dt_first": dts[0] if len(dts) else None,
                   "mass_rel_drift": abs(mass1 - mass0) / abs(mass0)},
    }
    if not args.no_small:
        log("1D shock tube and 2D vortex configurations")
        line["single_grid"] = configs_1d_2d()
    if args.recipes:
        log(f"every convection recipe on {args.recipes}^3")
        line["recipes"] = recipes_run(args.recipes, 3)
    if args.batched:
        shape = [int(x) for x in args.batched_shape.split(",")] if args.batched_shape else None
        line["batched"] = batched_run(args.batched, args.batched_steps, args.plain_sample,
                                      launch=1 if args.batched_launch == "graph" else 0,
                                      shape=shape)
    if args.served:
        line["served"] = served_run(args.served, args.served_rate, args.served_budget_gb,
                                    min(args.plain_sample, 24),
                                    launch=2 if args.served_launch == "auto" else 0,
                                    shape=[int(x) for x in args.served_shape.split(",")]
                                    if args.served_shape else None)
    if not args.no_cpu_baseline:
        log(f"CPU baseline: reference code on a {args.cpu_sample}^3 sample")
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
