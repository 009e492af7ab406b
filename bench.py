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
own translation units (oracle/_ref/libsfref.so, serial) on a bounded sample
of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mcell-updates/s FP64 WENO5+RK3 (% HBM peak); batched solves/s + p99 latency"
HBM_PEAK_GBPS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(n: int, steps: int):
    """The reference's own code (harness over its translation units) on an
    n^3 sample of the same workload, serial; falls back to the C
    restatement where the harness was not built."""
    from oracle import bind as ob
    from sonicsolver_amd import workloads as wl

    kw, q = wl.tgv_3d(n)
    desc = ob.make_desc(**kw)
    kind = "reference" if ob.ref_available() else "port"
    runner = ob.ref_run if kind == "reference" else ob.run
    t0 = time.perf_counter()
    runner(desc, q, steps, 0.4)
    dt = time.perf_counter() - t0
    what = ("reference translation units via oracle/_ref/libsfref.so" if kind == "reference"
            else "C restatement oracle/liboracle.so")
    return {
        "value": round(n ** 3 * steps / dt / 1e6, 4),
        "unit": "Mcell-updates/s",
        "cores": 1,
        "kind": kind,
        "sample": f"3D periodic TGV {n}^3, WENO5-JS/HLLC/SSPRK3 CFL 0.4, {steps} steps, "
                  f"{what}, serial, {dt:.1f} s",
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
    ap.add_argument("--cpu-sample", type=int, default=96, help="n of the n^3 CPU sample")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU only; the explicit hot path is measured "
              "on a single MI355X (multi-GPU sharding is a design note, not a measured "
              "configuration). Re-run with --gpus 1.", file=sys.stderr)
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
    kernel_names = {"sweep_x": "k_sweep_x (x faces + stage update + CFL fold)",
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
            "bound": "mfma",
            "pipe": "FP64 vector ALU (no dense contraction on this path); peak is MI355X's "
                    "FP64 figure, which is the same for vector and matrix FP64",
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
    if not args.no_cpu_baseline:
        log(f"CPU baseline: reference code on a {args.cpu_sample}^3 sample")
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
