"""Benchmark of the SPH step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--cpu-steps S]
                [--legs plain,batched,serving,reference,skewed] [--no-autotune]

Workload: the 158^3 = 3,944,312-particle column in a 6.32 m cube (317^3
cells), h = 0.02, dt = 1e-4, FP32 (synthetic workload 1), 200 timed steps by
default.  A step is one pass of the hot path (arrange, density, force,
integrate) over that state, resident in HBM.  Prints one JSON line; beside
the plain step it carries the batched (256 scenes), serving (Poisson
requests), skewed-occupancy (pool + fine sphere, 8.1 M particles) legs, the
reference's own default scene timed against its published frame time, and
the CPU baseline with a per-quantity parity check of workload 0 against the
FP64 oracle.
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
# rocprofv3 summaries of the plain workload (written with profsum.py)
TRAFFIC_PROFILE = ROOT / "profiles" / "hbm_traffic_plain_3.94M.md"
VALU_PROFILE = ROOT / "profiles" / "valu_plain_3.94M.md"


def profile_block(path):
    """The JSON block of a committed rocprof summary, or {} if absent."""
    if not path.exists():
        return {}
    import profsum
    return profsum.read_json_block(path)


def algorithmic_bytes(n, ncell, fused=False):
    """Bytes per launch that the step must move (reference layout, neighbour
    gathers cache-served): per particle 288 B, per cell 8 B.  With the
    integration fused into the force kernel, that kernel moves force's 44 B
    plus integrate's 100 B minus the 44 B integrate would read back (pos, ev,
    acc stay in registers): 100 B."""
    return {
        "arrange": 116 * n + 8 * ncell,
        "density": 28 * n,
        "force": (100 if fused else 44) * n,
        "integrate": 0 if fused else 100 * n,
        "step": 288 * n + 8 * ncell,
    }


def cpu_baseline(gsph, params, pos, steps, threads):
    """The CPU baseline (the oracle's FP32 OpenMP restatement of the reference
    step, timed on `threads` host threads over `steps` steps of the workload)
    and, with the oracle as the checker, the per-quantity errors of one GPU
    step of workload 0 against its FP64 step."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "tests"))
    import pyoracle as po
    from tolerance import step_errors

    op = po.params_from(params.as_dict())
    t0 = time.perf_counter()
    po.step_f32(op, pos, nsteps=steps, nthreads=threads)
    dt = time.perf_counter() - t0
    base = {
        "value": pos.shape[0] * steps / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "nproc": os.cpu_count(),
        "kind": "port",
        "sample": f"FP32 C++ restatement of the reference step (oracle/oracle.cpp, OpenMP on "
                  f"{threads} threads, the job's CPU share), {steps} steps of the same "
                  f"3,944,312-particle workload, {dt:.1f} s",
    }
    p0, pos0 = gsph.scene_config(0)
    n0 = pos0.shape[0]
    s = gsph.System(p0, pos0)
    s.step(1)
    got = s.read_state()
    s.close()
    op0 = po.params_from(p0.as_dict())
    ref = po.fp64_step(op0, pos0, np.zeros((n0, 3), np.float32), np.zeros((n0, 4), np.float32))
    errs = step_errors(*got, ref, op0)
    parity = {"workload": f"workload 0 (32^3 = {n0} particles), one step from rest",
              "bar": 1e-5, "max_error": errs,
              "measure": "each quantity relative to the scale its FP32 rounding follows "
                         "(tests/tolerance.py)"}
    return base, parity


def opt_dict(o):
    return {k: getattr(o, k) for k, _ in o._fields_}


def batched_leg(gsph, steps, warmup, count=256, loop_steps=5, no_autotune=False):
    import numpy as np

    params, _ = gsph.scene_config(0)
    worlds, blocks = gsph.batch_workload(count)
    counts = [len(b) for b in blocks]
    n = int(sum(counts))
    b = gsph.System(params, np.concatenate(blocks), batch_worlds=worlds, batch_counts=counts)
    tuned = b.options() if no_autotune else b.autotune(trial_steps=2)[0]
    b.enable_graph(True)
    b.step(warmup)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    wall = time.perf_counter() - t0
    batched_ps = n * steps / wall
    per_step_ms = wall * 1e3 / steps
    b.close()
    # the plain per-scene launch loop: one context (own stream) per scene
    systems = []
    for w, blk in zip(worlds, blocks):
        p, _ = gsph.scene_config(0)
        for a in range(3):
            p.world_size[a] = w[a]
        systems.append(gsph.System(p, blk))
    for s in systems:
        s.step(1)
    for s in systems:
        s.sync()
    t0 = time.perf_counter()
    for _ in range(loop_steps):
        for s in systems:
            s.step(1)
    for s in systems:
        s.sync()
    loop_wall = time.perf_counter() - t0
    loop_ps = n * loop_steps / loop_wall
    for s in systems:
        s.close()
    return {
        "workload": f"{count} dam-break scenes, N log-uniform 4096..262144 "
                    f"(total {n}), one context, hipGraph step",
        "scenes": count, "particles": n,
        "particle_steps_per_s": batched_ps, "scenes_per_s": count * steps / wall,
        "ms_per_step": per_step_ms,
        "plain_loop_particle_steps_per_s": loop_ps,
        "plain_loop_scenes_per_s": count * loop_steps / loop_wall,
        "speedup_vs_plain_loop": batched_ps / loop_ps,
        "autotuned_options": opt_dict(tuned),
    }


# The reference's own benchmark: assets/scene_default.json (3,944,312
# particles, h = 0.03, ~296 per cell, SMS path), mean frame 21.833 ms on an
# RTX 5090 (README.md:212, agent_docs/optimization_report.md:194).
REFERENCE_SCENE_MS = 21.833


def reference_scene_leg(gsph, steps=20, warmup=3, no_autotune=False):
    path = ROOT / "fast-general-gpu-sph-framework_amd" / "assets" / "scene_default.json"
    params, pos = gsph.scene_json(path)
    s = gsph.System(params, pos)
    # dense cells: the split is tuned too (the TRA path with pool masks can
    # beat the reference's SMS rule there)
    tuned = s.options() if no_autotune else s.autotune(trial_steps=2, tune_split=True)[0]
    s.step(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    st = s.step(5, timings=True)
    info = s.grid_info()
    s.close()
    n = pos.shape[0]
    # the reference's SMS build variants (CMakeLists.txt:59-100) under its own
    # split rule (every dense cell on the SMS path), same launch shape
    variants = {}
    v = gsph.System(params, pos)
    base = v.options()
    for tag, opts in (("register", {}), ("register_no_local_merge", {"sms_local_merge": 0}),
                      ("legacy", {"sms_legacy_staging": 1}),
                      ("register_same_cell_pair_force", {"same_cell_pair_force": 1}),
                      ("legacy_same_cell_pair_force", {"sms_legacy_staging": 1,
                                                       "same_cell_pair_force": 1})):
        o = gsph.Options.from_buffer_copy(base)
        for k, val in opts.items():
            setattr(o, k, val)
        v.set_options(o)
        v.step(3)
        v.sync()
        t0 = time.perf_counter()
        v.step(10)
        v.sync()
        vms = (time.perf_counter() - t0) * 1e3 / 10
        vst = v.step(3, timings=True)
        variants[tag] = {"ms_per_step": vms,
                         "stages_ms": {k: sum(x[k] for x in vst) / len(vst) for k in vst[0]}}
    v.close()
    return {
        "workload": "the reference's default scene (assets/scene_default.json): 3,944,312 "
                    "particles, h=0.03, ~296 per cell",
        "ms_per_step": ms, "particle_steps_per_s": n * 1e3 / ms,
        "stages_ms": {k: sum(x[k] for x in st) / len(st) for k in st[0]},
        "sms_tasks": info["tasks"], "tra_particles": info["tra"],
        "path": ("SMS tasks for every dense cell" if info["tra"] == 0 else
                 "TRA for every cell (long rows keep their pair masks in the pool)"
                 if info["tasks"] == 0 else "mixed TRA / SMS"),
        "path_reason": "the autotuner times the reference split rule, tuned thresholds and "
                       "all-TRA and keeps the fastest",
        "reference_published_ms": REFERENCE_SCENE_MS,
        "reference_published_hw": "RTX 5090 (README.md:212)",
        "speedup_vs_reference_published": REFERENCE_SCENE_MS / ms,
        "autotuned_options": opt_dict(tuned),
        "sms_variants_reference_split": variants,
    }


def skewed_leg(gsph, steps=100, warmup=5, no_autotune=False):
    """Workload 2: a 0.01-spaced pool (256 x 64 x 256) plus a 0.005-spaced
    sphere dropped from 1 m above (8.1 M particles, 0 to 64 per cell): the
    configuration that exercises the TRA/SMS split.  Autotuned including the
    split; per-stage times and the chosen split."""
    params, pos = gsph.scene_config(2)
    s = gsph.System(params, pos)
    tuned, tuned_ms = (s.options(), None) if no_autotune else s.autotune(trial_steps=2,
                                                                           tune_split=True)
    s.step(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    st = s.step(10, timings=True)
    info = s.grid_info()
    s.close()
    n = pos.shape[0]
    return {
        "workload": f"pool 256x64x256 at 0.01 + sphere at 0.005 ({n} particles), h=0.02, "
                    f"{steps} timed steps",
        "particles": n, "ms_per_step": ms, "particle_steps_per_s": n * 1e3 / ms,
        "stages_ms": {k: sum(x[k] for x in st) / len(st) for k in st[0]},
        "split": {"split_mode": tuned.split_mode, "sms_min_self": tuned.sms_min_self,
                  "sms_min_total": tuned.sms_min_total, "tra_particles": info["tra"],
                  "sms_tasks": info["tasks"]},
        "autotuned_options": opt_dict(tuned),
    }


def serving_leg(gsph, requests=96, seed=20261207, budget=200e9, load=0.7):
    """Workload 4: requests with N log-uniform in [8192, 1048576] and steps
    uniform in 50..500, 200 GB HBM budget.  Three runs of the same requests:
    (1) served capacity at saturation (every request queued at t = 0); (2)
    latency under Poisson arrivals whose offered load is `load` x that
    capacity; (3) the plain loop: each request alone in its own context, one
    after the other, its final state read back as the served results are."""
    import numpy as np

    worlds, blocks, steps, rng = gsph.serving_workload(requests, seed)
    ns = np.array([len(b) for b in blocks], np.float64)
    work = float(np.sum(ns * np.array(steps)))
    params, _ = gsph.scene_config(0)

    def serve(arrivals):
        srv = gsph.Server(params, hbm_budget_bytes=budget, max_scenes=4096)
        for w, blk, st, a in zip(worlds, blocks, steps, arrivals):
            srv.submit(w, blk, st, float(a))
        srv.run(realtime=True)
        st = srv.stats()
        srv.close()
        return st

    sat = serve(np.zeros(requests))
    capacity = sat["particle_steps"] / sat["wall"]
    print(f"[bench] serving capacity {capacity / 1e9:.2f} G particle-steps/s", file=sys.stderr,
          flush=True)
    lam = load * capacity / (work / requests)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, requests))
    lat = serve(arrivals)
    # plain loop: one plain context per request, back to back, each result
    # (pos, vel, ev, the state the server returns) read back into pinned host
    # buffers allocated beforehand, as the server's are at submission
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
    hip.hipHostFree.argtypes = [ctypes.c_void_p]
    outs = []
    for b in blocks:
        ptrs = []
        for width in (16, 12, 16):
            q = ctypes.c_void_p()
            if hip.hipHostMalloc(ctypes.byref(q), len(b) * width, 0) != 0:
                raise RuntimeError("hipHostMalloc failed")
            ptrs.append(q)
        outs.append(ptrs)
    t0 = time.perf_counter()
    for w, blk, st, (op, ov, oe) in zip(worlds, blocks, steps, outs):
        p, _ = gsph.scene_config(0)
        for a in range(3):
            p.world_size[a] = w[a]
        s = gsph.System(p, blk)
        s.step(st)
        gsph._check(s._lib.gsph_read_state(s.h, op, ov, oe, None), s.h)
        s.close()
    loop_ps = work / (time.perf_counter() - t0)
    for ptrs in outs:
        for q in ptrs:
            hip.hipHostFree(q)
    return {
        "workload": f"{requests} requests, N log-uniform 8192..1048576 (total {int(ns.sum())}), "
                    f"steps 50..500, {budget / 1e9:.0f} GB HBM budget",
        "capacity_particle_steps_per_s": capacity,
        "capacity_wall_s": sat["wall"], "capacity_recompositions": sat["recompositions"],
        "capacity_max_resident": sat["max_active"], "capacity_peak_bytes": sat["peak_bytes"],
        "plain_loop_particle_steps_per_s": loop_ps,
        "capacity_vs_plain_loop": capacity / loop_ps,
        "load": load, "poisson_rate_per_s": lam,
        "completed": lat["completed"], "failed": lat["failed"], "rejected": lat["rejected"],
        "p50_latency_s": lat["p50"], "p99_latency_s": lat["p99"],
        "particle_steps_per_s": lat["particle_steps"] / lat["wall"], "wall_s": lat["wall"],
        "max_resident": lat["max_active"], "recompositions": lat["recompositions"],
        "engine_steps": lat["engine_steps"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--legs", default="plain,batched,serving,reference,skewed")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default options (profiling runs: no trial launches)")
    ap.add_argument("--probe-copy", type=float, default=0.0,
                    help="only run a plain copy kernel over this many bytes (counter calibration)")
    ap.add_argument("--probe-valu", type=int, default=0,
                    help="only run the FP32 FMA-chain kernel for this many iterations (VALU peak)")
    args = ap.parse_args()
    if args.probe_valu > 0:
        import gsph
        ms, ops = gsph.probe_valu(args.probe_valu)
        print(json.dumps({"probe_valu_iters": args.probe_valu, "ms": ms, "lane_ops": ops,
                          "tlane_ops_per_s": ops / (ms * 1e-3) / 1e12}))
        return
    if args.probe_copy > 0:
        import gsph
        ms = gsph.probe_copy(args.probe_copy, 5)
        print(json.dumps({"probe_copy_bytes": args.probe_copy, "avg_ms": ms,
                          "gbs": 2 * args.probe_copy / (ms * 1e-3) / 1e9}))
        return
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
    if args.no_autotune:
        tuned, tuned_ms = s.options(), None
    else:
        tuned, tuned_ms = s.autotune(trial_steps=3)   # launch shape only: same bits
    print(f"[bench] options: {opt_dict(tuned)} (autotune {tuned_ms} ms/step)", file=sys.stderr,
          flush=True)
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
    # the VALU peak of this GPU: independent FP32 FMA chains on 8 waves per SIMD
    probe_ms, probe_ops = gsph.probe_valu(4000)
    valu_peak = probe_ops / (probe_ms * 1e-3) / 1e12   # T lane-ops/s

    kern = {k: {"avg_ms": v[0] / max(v[1], 1), "launches": v[1]} for k, v in kt.items()}
    fused = "integrate" not in kern   # the full step integrates inside the force kernel
    alg = algorithmic_bytes(n, ncell, fused)
    dom = max((k for k in ("density", "force", "integrate") if k in kern),
              key=lambda k: kern[k]["avg_ms"])
    dom_ms = kern[dom]["avg_ms"]
    achieved = alg[dom] / (dom_ms * 1e-3) / 1e9
    stage_avg = {k: float(np.mean([x[k] for x in stages])) for k in stages[0]}
    stage_gbs = {k: alg[k] / (stage_avg[k] * 1e-3) / 1e9
                 for k in ("arrange", "density", "force", "integrate")
                 if alg[k] > 0 and stage_avg[k] > 0}
    step_gbs = alg["step"] / (ms_per_step * 1e-3) / 1e9

    def single_scene(table, stage):
        """A stage's single-scene kernels summed per launch: its lean pass
        (k_<stage>_lean<false...>) and its general pass (k_<stage><false...>),
        any staging-depth argument; the lanes-active share recomputed from
        the summed counters."""
        parts = [v for k, v in table.items()
                 if k.startswith(f"k_{stage}<false") or k.startswith(f"k_{stage}_lean<false")]
        if not parts:
            return {}
        out = {}
        for key in ("bytes", "SQ_THREAD_CYCLES_VALU", "SQ_INSTS_VALU"):
            vals = [p[key] for p in parts if key in p]
            if vals:
                out[key] = sum(vals)
        if out.get("SQ_INSTS_VALU"):
            out["valu_util_pct"] = 100.0 * out["SQ_THREAD_CYCLES_VALU"] / (64.0 * out["SQ_INSTS_VALU"])
        return out

    traffic = profile_block(TRAFFIC_PROFILE).get("traffic", {})
    valu = profile_block(VALU_PROFILE).get("counters", {})
    dom_traffic = single_scene(traffic, dom).get("bytes")
    # active-lane VALU operations of one launch of the dominant kernel
    # (SQ_THREAD_CYCLES_VALU of the committed counter profile) over its live
    # duration, against the FMA-chain peak measured above
    dom_lane_ops = single_scene(valu, dom).get("SQ_THREAD_CYCLES_VALU")
    valu_achieved = dom_lane_ops / (dom_ms * 1e-3) / 1e12 if dom_lane_ops else None
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
        # "bound" is the roofline this line is priced against (HBM: the
        # path has no dense contraction for MFMA); "limiter" is what the
        # counters show actually bounds the kernel, the larger fraction
        "roofline": {"bound": "hbm", "limiter": ("valu" if valu_achieved and
                                                 valu_achieved / valu_peak > achieved / HBM_PEAK_GBS
                                                 else "hbm"),
                     "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": dom_traffic, "algorithmic_bytes": alg[dom], "avg_ms": dom_ms,
                     "traffic_source": TRAFFIC_PROFILE.name if dom_traffic else None,
                     "valu": {"lane_ops": dom_lane_ops, "achieved": valu_achieved,
                              "peak": valu_peak, "unit": "T lane-ops/s",
                              "frac": valu_achieved / valu_peak if valu_achieved else None,
                              "lanes_active_pct": single_scene(valu, dom).get("valu_util_pct"),
                              "source": VALU_PROFILE.name if dom_lane_ops else None}},
        "step_roofline": {"algorithmic_bytes": alg["step"], "achieved_gbs": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS},
        "stages_ms": stage_avg,
        "stages_gbs": stage_gbs,
        "kernels_ms": {k: v["avg_ms"] for k, v in kern.items()},
        "split": {"tra": info["tra"], "sms_tasks": info["tasks"]},
        "autotuned_options": opt_dict(tuned),
        "checksum": [float(x) for x in cs],
    }
    s.close()
    legs = args.legs.split(",")
    if "batched" in legs:
        print("[bench] batched leg", file=sys.stderr, flush=True)
        out["batched"] = batched_leg(gsph, args.batched_steps, 3, no_autotune=args.no_autotune)
        print(f"[bench] batched {out['batched']['speedup_vs_plain_loop']:.2f}x plain loop",
              file=sys.stderr, flush=True)
    if "serving" in legs:
        print("[bench] serving leg", file=sys.stderr, flush=True)
        out["serving"] = serving_leg(gsph)
        print(f"[bench] serving capacity {out['serving']['capacity_vs_plain_loop']:.2f}x the plain "
              f"loop; p99 {out['serving']['p99_latency_s']:.3f} s at 0.7 load", file=sys.stderr,
              flush=True)
    if "reference" in legs:
        print("[bench] reference scene leg", file=sys.stderr, flush=True)
        out["reference_scene"] = reference_scene_leg(gsph, no_autotune=args.no_autotune)
        print(f"[bench] reference scene {out['reference_scene']['ms_per_step']:.2f} ms/step",
              file=sys.stderr, flush=True)
    if "skewed" in legs:
        print("[bench] skewed-occupancy leg", file=sys.stderr, flush=True)
        out["skewed"] = skewed_leg(gsph, no_autotune=args.no_autotune)
        print(f"[bench] skewed {out['skewed']['ms_per_step']:.2f} ms/step", file=sys.stderr,
              flush=True)
    if not args.no_cpu:
        # the job's CPU share (OMP_NUM_THREADS is set to it on the GPU box;
        # nproc may show the whole host)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        print(f"[bench] cpu baseline {args.cpu_steps} steps on {threads} threads",
              file=sys.stderr, flush=True)
        out["cpu_baseline"], out["parity"] = cpu_baseline(gsph, params, pos, args.cpu_steps,
                                                          threads)
    else:
        out["cpu_baseline"] = None
        out["parity"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
