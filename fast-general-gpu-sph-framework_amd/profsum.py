"""Summaries of rocprofv3 CSV output for the SPH step kernels.

    python profsum.py stats    <kernel_stats.csv>
    python profsum.py counters <counter_collection.csv> [...]
    python profsum.py traffic  --fetch F.csv --write W.csv --probe-fetch PF.csv \
                               --probe-write PW.csv --probe-bytes B

`stats` turns a `--kernel-trace --stats` table into per-kernel average
durations; `counters` averages every counter per dispatch of each kernel and
adds the active-lane VALU operations per CU-cycle, their share of the FMA-chain
probe's rate (`k_fma_probe`, run by bench.py in the same profile) and the
lanes active per VALU instruction, when their inputs were collected; `traffic`
converts FETCH_SIZE / WRITE_SIZE (KiB per dispatch) into HBM bytes per launch,
each scaled by what the same counter reads for the plain copy kernel `k_copy`
over a buffer of known size (bench.py --probe-copy B).  Output is Markdown
with one fenced JSON block that bench.py reads back.  --only k_a,k_b keeps
the kernels whose names start with one of the given prefixes (the probe
kernel of `counters` is always used for the peak).
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sys

CU_NUM = 256


def short_name(kernel: str) -> str:
    """k_name<template args> of a demangled kernel name."""
    m = re.search(r"(k_\w+(?:<[^>]*>)?)", kernel)
    return m.group(1) if m else kernel.split("(")[0][:40]


def per_dispatch(path: str) -> dict[str, dict[str, float]]:
    """{kernel: {counter: mean over dispatches}} of one counter_collection.csv."""
    agg: dict = collections.defaultdict(lambda: collections.defaultdict(float))
    disp: dict = collections.defaultdict(set)
    with open(path) as f:
        for r in csv.DictReader(f):
            k = short_name(r["Kernel_Name"])
            agg[k][r["Counter_Name"]] += float(r["Counter_Value"])
            disp[k].add(r["Dispatch_Id"])
    out = {}
    for k, d in agg.items():
        n = len(disp[k])
        out[k] = {c: v / n for c, v in d.items()}
        out[k]["dispatches"] = n
    return out


def keep(args, kernel: str) -> bool:
    return not args.only or any(kernel.startswith(p) for p in args.only.split(","))


def cmd_stats(args):
    rows = list(csv.DictReader(open(args.csv)))
    print("| kernel | calls | avg (us) | min (us) | max (us) | share of kernel time |")
    print("|---|---:|---:|---:|---:|---:|")
    data = {}
    for r in rows:
        k = short_name(r["Name"])
        if not keep(args, k):
            continue
        calls = int(r["Calls"])
        avg = float(r["AverageNs"]) / 1e3
        data[k] = {"calls": calls, "avg_us": avg}
        print(f"| {k} | {calls} | {avg:.2f} | {float(r['MinNs']) / 1e3:.2f} | "
              f"{float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.1f}% |")
    print("\n```json\n" + json.dumps({"kernel_avg_us": data}, indent=1) + "\n```")


def cmd_counters(args):
    merged: dict = collections.defaultdict(dict)
    for p in args.csv:
        for k, d in per_dispatch(p).items():
            merged[k].update(d)
    names = sorted({c for d in merged.values() for c in d if c != "dispatches"})

    def lanes_per_cu_cycle(d):
        # active-lane VALU operations per CU per cycle; GRBM_GUI_ACTIVE is
        # summed over the 8 XCDs, so one XCD's count is the kernel's cycles
        if "SQ_THREAD_CYCLES_VALU" in d and d.get("GRBM_GUI_ACTIVE"):
            return d["SQ_THREAD_CYCLES_VALU"] / CU_NUM / (d["GRBM_GUI_ACTIVE"] / 8)
        return None

    # the FMA-chain probe (k_fma_probe) in the same run is the peak
    peak = next((lanes_per_cu_cycle(d) for k, d in merged.items() if k.startswith("k_fma_probe")),
                None)
    print("| kernel | dispatches | " + " | ".join(names) +
          " | VALU lane-ops per CU-cycle | of the FMA-chain peak | lanes active per VALU instruction |")
    print("|---|---:|" + "---:|" * len(names) + "---:|---:|---:|")
    out = {}
    for k in sorted(merged):
        if not keep(args, k) and not k.startswith("k_fma_probe"):
            continue
        d = merged[k]
        lanes = lanes_per_cu_cycle(d)
        frac = lanes / peak if lanes is not None and peak else None
        util = None
        if "SQ_THREAD_CYCLES_VALU" in d and d.get("SQ_INSTS_VALU"):
            util = 100.0 * d["SQ_THREAD_CYCLES_VALU"] / (d["SQ_INSTS_VALU"] * 64)
        out[k] = dict(d, valu_lanes_per_cu_cycle=lanes, valu_frac_of_peak=frac, valu_util_pct=util)
        cells = " | ".join(f"{d.get(c, float('nan')):.4g}" for c in names)
        print(f"| {k} | {int(d['dispatches'])} | {cells} | "
              f"{'' if lanes is None else f'{lanes:.1f}'} | "
              f"{'' if frac is None else f'{100 * frac:.1f}%'} | "
              f"{'' if util is None else f'{util:.1f}%'} |")
    print("\n```json\n" + json.dumps({"counters": out}, indent=1) + "\n```")


def cmd_traffic(args):
    fetch = per_dispatch(args.fetch)
    write = per_dispatch(args.write)
    pf = per_dispatch(args.probe_fetch)["k_copy"]["FETCH_SIZE"]
    pw = per_dispatch(args.probe_write)["k_copy"]["WRITE_SIZE"]
    # the copy kernel reads and writes probe_bytes each
    f_scale = args.probe_bytes / (pf * 1024.0)
    w_scale = args.probe_bytes / (pw * 1024.0)
    print(f"Copy-kernel calibration over {args.probe_bytes:.0f} B: FETCH_SIZE reads "
          f"{pf:.6g} KiB (scale {f_scale:.4f}), WRITE_SIZE reads {pw:.6g} KiB "
          f"(scale {w_scale:.4f}).\n")
    print("| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM read B | HBM write B | HBM total B |")
    print("|---|---:|---:|---:|---:|---:|")
    out = {}
    for k in sorted(set(fetch) & set(write)):
        if not keep(args, k):
            continue
        fk, wk = fetch[k]["FETCH_SIZE"], write[k]["WRITE_SIZE"]
        rb, wb = fk * 1024 * f_scale, wk * 1024 * w_scale
        out[k] = {"fetch_kib": fk, "write_kib": wk, "read_bytes": rb, "write_bytes": wb,
                  "bytes": rb + wb}
        print(f"| {k} | {fk:.6g} | {wk:.6g} | {rb:.4g} | {wb:.4g} | {rb + wb:.4g} |")
    print("\n```json\n" + json.dumps({"fetch_scale": f_scale, "write_scale": w_scale,
                                      "probe_bytes": args.probe_bytes, "traffic": out},
                                     indent=1) + "\n```")


def read_json_block(path) -> dict:
    """The fenced JSON block of a summary written by this module."""
    text = open(path).read()
    m = re.search(r"```json\n(.*?)\n```", text, re.S)
    return json.loads(m.group(1)) if m else {}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    c = sub.add_parser("counters")
    c.add_argument("csv", nargs="+")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--probe-fetch", required=True)
    t.add_argument("--probe-write", required=True)
    t.add_argument("--probe-bytes", type=float, required=True)
    for p in (s, c, t):
        p.add_argument("--only", default="", help="comma-separated kernel name prefixes")
    args = ap.parse_args(argv)
    {"stats": cmd_stats, "counters": cmd_counters, "traffic": cmd_traffic}[args.cmd](args)


if __name__ == "__main__":
    sys.exit(main())
