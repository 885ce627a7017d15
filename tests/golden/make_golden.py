"""Writes the golden fixtures in tests/golden/*.npz from the CPU oracle.

Each fixture holds the inputs (params, pos4, vel3, ev4) and the oracle's
outputs: the arrange arrays (bit-exact targets) and one FP64 step (rho,
pressure, acceleration, its tolerance scale, integrated state).  The oracle
itself is pinned by tests/test_oracle_cpu.py; the fixtures freeze its output
so that a later change to it, or a GPU run on a box without the oracle build,
is still checked against the same numbers.

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path[:0] = [str(HERE.parent), str(HERE.parent.parent / "oracle")]

import pyoracle as po  # noqa: E402
from helpers import make_params  # noqa: E402


def cell_cluster(centre_cell, n, h, rng):
    lo = np.array(centre_cell, np.float64) * h
    pts = np.zeros((n, 4), np.float32)
    pts[:, :3] = lo + rng.uniform(0.02 * h, 0.98 * h, (n, 3))
    return pts


def fixtures():
    rng = np.random.default_rng(20261203)
    h = 0.02
    out = {}
    p = make_params(h=h, world=(0.2, 0.2, 0.2))
    out["tiny_pair"] = (p, np.array([[0.1, 0.1, 0.1, 0], [0.113, 0.104, 0.097, 0]], np.float32))
    out["lattice_4"] = (p, po.scene_block([0.07, 0.07, 0.07], 4, 4, 4, 0.01, 0.0, 1))
    out["jitter_8"] = (p, po.scene_block([0.06, 0.06, 0.06], 8, 8, 8, 0.01, 0.01, 2))
    p2 = make_params(h=h, world=(0.24, 0.24, 0.24))
    for n in (14, 15, 33, 36, 37):
        out[f"dense_cell_{n}"] = (p2, cell_cluster((5, 5, 5), n, h, rng))
    out["odd_tasks"] = (p2, np.concatenate([cell_cluster((3, 3, 3), 70, h, rng),
                                            cell_cluster((7, 4, 6), 20, h, rng)]))
    # same-cell pairs whose x bounds meet (compact) and do not (widened)
    rows = []
    for j in range(64):
        rows.append([0.08 + 0.0075, 0.08 + 0.0025 + 0.005 * (j % 4), 0.08 + 0.005 * (j // 16) + 0.0025, 0])
    out["is_same_compact"] = (p2, np.array(rows, np.float32))
    out["is_same_widened"] = (p2, cell_cluster((4, 4, 4), 64, h, rng))
    return out


def record(p, pos):
    n = len(pos)
    g = po.arrange(p, pos)
    r = po.fp64_step(p, pos, np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32))
    params = np.array([p.kernel_h, p.mass, p.rest_density, p.gas_constant, p.viscosity,
                       p.time_step, *p.gravity, p.wall_damping, p.surface_normal,
                       p.surface_coe, *p.world_size], np.float32)
    return dict(params=params, pos=pos, hash=g["hash"], perm=g["perm"], begin=g["begin"],
                count=g["count"], req=g["req"], compact=g["compact"],
                middle_raw=np.int32(g["middle_raw"]), tasks_count=np.int32(g["tasks_count"]),
                tasks=g["tasks"].view(np.uint8).reshape(-1, 32) if len(g["tasks"]) else
                np.zeros((0, 32), np.uint8),
                rho=r["rho"], pressure=r["pressure"], acc=r["acc"], acc_scale=r["acc_scale"],
                pos1=r["pos"], vel1=r["vel"], ev1=r["ev"])


def main():
    for name, (p, pos) in fixtures().items():
        np.savez_compressed(HERE / f"{name}.npz", **record(p, pos))
        print("wrote", name, len(pos))


if __name__ == "__main__":
    main()
