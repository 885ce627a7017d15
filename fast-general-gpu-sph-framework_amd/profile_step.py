"""Runs a few steps of one named workload with fixed options, for profiling
the step kernels under rocprofv3 (kernel trace, counters) without autotune
trial launches or other workloads mixed into the same profile.

    python profile_step.py WORKLOAD [--steps K] [--warmup W] [--split reference|all_tra]
                           [--legacy] [--same-cell]

Workloads:
  plain      158^3 = 3,944,312-particle column in a 6.32 m cube (317^3 cells)
  reference  the reference's default scene (assets/scene_default.json):
             3,944,312 particles, h = 0.03, about 296 per cell
  skewed     pool 256 x 64 x 256 at 0.01 plus a sphere at 0.005 (8,137,432)
  batched    256 dam-break scenes, N log-uniform 4096..262144 (17,280,876),
             one context, hipGraph step

--split reference keeps the reference split rule (dense cells on the SMS
path, the default); all_tra puts every cell on the TRA path (long rows keep
their accepted-pair masks in the pool).  --legacy and --same-cell select the
reference's SMS build variants.  Prints one JSON line: the workload, the
options, the split and the per-stage times of the last steps.
"""
from __future__ import annotations

import argparse
import json
import pathlib
import sys

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))

import gsph  # noqa: E402

ALL_TRA = {"split_mode": 1, "sms_min_self": 1 << 30, "sms_min_total": 1 << 30}


def build(workload: str):
    import numpy as np
    if workload == "plain":
        return gsph.scene_config(1) + (None, None)
    if workload == "skewed":
        return gsph.scene_config(2) + (None, None)
    if workload == "reference":
        return gsph.scene_json(HERE / "assets" / "scene_default.json") + (None, None)
    if workload == "batched":
        params, _ = gsph.scene_config(0)
        worlds, blocks = gsph.batch_workload(256)
        return params, np.concatenate(blocks), worlds, [len(b) for b in blocks]
    raise SystemExit(f"unknown workload {workload}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["plain", "reference", "skewed", "batched"])
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--split", choices=["reference", "all_tra"], default="reference")
    ap.add_argument("--legacy", action="store_true")
    ap.add_argument("--same-cell", action="store_true")
    args = ap.parse_args()
    params, pos, worlds, counts = build(args.workload)
    o = gsph.options_default()
    if args.split == "all_tra":
        for k, v in ALL_TRA.items():
            setattr(o, k, v)
    o.sms_legacy_staging = int(args.legacy)
    o.same_cell_pair_force = int(args.same_cell)
    if worlds is None:
        s = gsph.System(params, pos, options=o)
    else:
        s = gsph.System(params, pos, options=o, batch_worlds=worlds, batch_counts=counts)
        s.enable_graph(True)
    print(f"[profile_step] {args.workload} n={len(pos)} split={args.split}", file=sys.stderr,
          flush=True)
    s.step(args.warmup)
    s.sync()
    st = s.step(args.steps, timings=worlds is None)
    s.sync()
    info = s.grid_info()
    out = {"workload": args.workload, "particles": int(len(pos)), "steps": args.steps,
           "split": {"tra": info["tra"], "sms_tasks": info["tasks"]},
           "options": {k: getattr(o, k) for k, _ in gsph.Options._fields_}}
    if st:
        out["stages_ms"] = {k: sum(x[k] for x in st) / len(st) for k in st[0]}
    s.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
