"""Benchmark of the SPH step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--cpu-steps S]

Workload: the 158^3 = 3,944,312-particle column in a 6.32 m cube (317^3
cells), h = 0.02, dt = 1e-4, FP32 (synthetic workload 1).  A step is one
pass of the hot path (arrange, density, force, integrate) over that state,
resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "fast-general-gpu-sph-framework_amd"))

METRIC = "particle-steps/s at 3.94M; batched scenes/s + p99 latency vs plain"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(n, ncell):
    """Bytes per launch that the step must move (reference layout, neighbour
    gathers cache-served): per particle 288 B, per cell 8 B."""
    return {
        "arrange": 116 * n + 8 * ncell,
        "density": 28 * n,
        "force": 44 * n,
        "integrate": 100 * n,
        "step": 288 * n + 8 * ncell,
    }


def cpu_baseline(params, pos, steps, threads):
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po

    op = po.params_from(params.as_dict())
    t0 = time.perf_counter()
    po.step_f32(op, pos, nsteps=steps, nthreads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": pos.shape[0] * steps / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP32 C++ restatement of the reference step (oracle/oracle.cpp, OpenMP), "
                  f"{steps} steps of the same 3,944,312-particle workload, {dt:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--config", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); "
              "the path shards by space across GPUs by design, but no multi-GPU "
              "run is part of this build.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import gsph

    params, pos = gsph.scene_config(args.config)
    n = pos.shape[0]
    c = gsph.derive_constants(params)
    ncell = c.grid_size[0] * c.grid_size[1] * c.grid_size[2]
    print(f"[bench] workload {args.config}: n={n} cells={ncell}", file=sys.stderr, flush=True)
    s = gsph.System(params, pos)
    s.step(args.warmup)
    s.sync()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    # timed region 1: back-to-back steps, wall clock
    s.sync()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = n * args.steps / wall
    print(f"[bench] {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)

    # timed region 2: the same steps with HIP events around every kernel on
    # the context's stream -> per-kernel average durations
    s.kernel_timing(True)
    s.step(args.steps)
    kt = s.kernel_times()
    s.kernel_timing(False)
    # timed region 3: per-stage events (the reference's warm-up timing)
    stages = s.step(min(args.steps, 20), timings=True)
    s.sync()
    info = s.grid_info()
    cs = s.checksum()

    alg = algorithmic_bytes(n, ncell)
    kern = {k: {"avg_ms": v[0] / max(v[1], 1), "launches": v[1]} for k, v in kt.items()}
    dom = max(("density", "force", "integrate"), key=lambda k: kern[k]["avg_ms"])
    dom_ms = kern[dom]["avg_ms"]
    achieved = alg[dom] / (dom_ms * 1e-3) / 1e9
    stage_avg = {k: float(np.mean([x[k] for x in stages])) for k in stages[0]}
    stage_gbs = {k: alg[k] / (stage_avg[k] * 1e-3) / 1e9
                 for k in ("arrange", "density", "force", "integrate")}
    step_gbs = alg["step"] / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain single scene: 158^3 = 3,944,312-particle column in a "
                               "6.32 m cube (317^3 cells), h=0.02, dt=1e-4, FP32",
                   "particles": n, "cells": ncell, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes": alg[dom], "avg_ms": dom_ms},
        "step_roofline": {"algorithmic_bytes": alg["step"], "achieved_gbs": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS},
        "stages_ms": stage_avg,
        "stages_gbs": stage_gbs,
        "kernels_ms": {k: v["avg_ms"] for k, v in kern.items()},
        "split": {"tra": info["tra"], "sms_tasks": info["tasks"]},
        "checksum": [float(x) for x in cs],
    }
    s.close()
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        print(f"[bench] cpu baseline {args.cpu_steps} steps on {threads} threads",
              file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(params, pos, args.cpu_steps, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
