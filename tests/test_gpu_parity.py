"""GPU parity tests: the HIP path through the C ABI against the CPU oracle.

Bars:
  * arrange outputs (sorted micro hash, stable permutation, cell begin/count,
    split req, compact index, middle value, SMS tasks and pairing) are
    bit-exact;
  * density (1/rho), pressure, acceleration, integrated position and velocity
    after one step are within REL_TOL = 1e-5 of the FP64 oracle, each relative
    to its own scale (see tolerance.py);
  * the state checksum after 10 steps agrees with the FP32 CPU reference step
    within CHECKSUM_TOL.
"""
import numpy as np
import pytest

from helpers import (cfg0_params, dense_scene, make_params, mixed_scene, skewed_scene,
                     to_gsph_params)
from tolerance import REL_TOL, step_errors

pytestmark = pytest.mark.gpu

CHECKSUM_TOL = 2e-6
TASK_FIELDS = ("is_same", "paired_bounds", "cellid", "cell_pos", "p_offset", "xxi", "xxx",
               "yyi", "yyy", "zzi", "zzz", "cell_begin", "cell_particle_count")


def banner(msg):
    print(f"\n[gpu-case] {msg}", flush=True)


def scene(po, name):
    if name == "cfg0":
        return cfg0_params(po)
    if name == "dense":
        return dense_scene(po)
    if name == "mixed":
        return mixed_scene(po)
    if name == "dense64":
        # cells of exactly 64 (lattice aligned with the grid): every task
        # pair is a full isSame pair of one cell
        p = make_params(h=0.02, world=(0.24, 0.24, 0.24))
        return p, po.scene_block([0.06, 0.06, 0.06], 16, 16, 16, 0.005, 0.01, 1)
    if name == "odd_tasks":
        # one cell row of dense cells whose task count is odd
        # a dense block (20 tasks) plus one isolated cell of 20 particles
        # (req (20 + 27) >> 5 = 1): 21 tasks, so judgeTask pads the tail
        p = make_params(h=0.02, world=(0.2, 0.2, 0.2))
        blk = po.scene_block([0.06, 0.06, 0.06], 9, 8, 8, 0.005, 0.01, 9)
        rng = np.random.default_rng(9)
        extra = np.zeros((20, 4), np.float32)
        extra[:, :3] = rng.uniform(0.1405, 0.1595, (20, 3))
        return p, np.concatenate([blk, extra])
    if name == "sparse_random":
        rng = np.random.default_rng(4)
        p = make_params(h=0.02, world=(0.5, 0.3, 0.4))
        pos = np.zeros((5000, 4), np.float32)
        pos[:, :3] = rng.uniform([0.031] * 3, [0.469, 0.269, 0.369], (5000, 3))
        return p, pos
    raise KeyError(name)


SCENES = ["cfg0", "dense", "dense64", "mixed", "odd_tasks", "sparse_random"]


def make_system(gsph, op, pos, **opts):
    o = gsph.options_default()
    for k, v in opts.items():
        setattr(o, k, v)
    return gsph.System(to_gsph_params(gsph, op), pos, options=o)


def compare_grid(g, ref):
    for k in ("hash", "perm", "begin", "count", "req", "compact"):
        assert np.array_equal(g[k], ref[k]), k
    assert g["middle_raw"] == ref["middle_raw"]
    assert g["tasks"] is not None and len(g["tasks"]) == len(ref["tasks"])
    T = ref["tasks_count"]
    assert g["tasks"].shape[0] == T + (T & 1)
    # every field of every descriptor, the odd-tail pad included
    for f in TASK_FIELDS:
        assert np.array_equal(g["tasks"][f], ref["tasks"][f]), f
    tra = ref["middle_raw"] if ref["middle_raw"] >= 0 else len(ref["perm"])
    assert g["tra"] == tra


def compare_pair_stats(s, ref):
    """All seven getSmsTaskPairStats fields against the oracle's."""
    exp = po_pair_stats(ref)
    got = s.pair_stats()
    assert got == exp, (got, exp)
    return got


def po_pair_stats(ref):
    import pyoracle
    return pyoracle.pair_stats(ref["tasks"], ref["tasks_count"], ref["count"])


@pytest.mark.parametrize("name", SCENES)
def test_arrange_bit_exact(gsph, po, name):
    op, pos = scene(po, name)
    banner(f"arrange {name} n={len(pos)}")
    s = make_system(gsph, op, pos)
    s.arrange()
    g = s.read_grid()
    ref = po.arrange(op, pos)
    compare_grid(g, ref)
    if name in ("dense", "dense64", "mixed", "odd_tasks"):
        assert ref["tasks_count"] > 0
    st = compare_pair_stats(s, ref)
    print(st, flush=True)
    if name == "dense64":
        assert st["same_cell_pairs"] == ref["tasks_count"] // 2
        assert st["full_is_same_pairs"] == st["is_same_pairs"] > 0
    if name == "odd_tasks":
        assert ref["tasks_count"] & 1
        assert st["partial_is_same_pairs"] > 0
    s.close()


def test_split_switches_between_arranges(gsph, po):
    """SMS work appears, disappears and appears again in one context (the
    split rule switched between arranges): each arrange's grid is the
    oracle's, and without SMS tasks the compact index is the identity, every
    req zero, no task and no middle."""
    op, pos = scene(po, "mixed")
    n = len(pos)
    ref = po.arrange(op, pos)
    assert ref["tasks_count"] > 0
    tra_only = dict(ref)
    tra_only.update(compact=np.arange(n, dtype=ref["compact"].dtype), req=np.zeros_like(ref["req"]),
                    middle_raw=-1, tasks_count=0, tasks=ref["tasks"][:0])
    banner(f"split switches n={n}")
    s = make_system(gsph, op, pos)
    for k, opts in enumerate(({}, ALL_TRA, {}, ALL_TRA, {})):
        o = gsph.options_default()
        for key, v in opts.items():
            setattr(o, key, v)
        s.set_options(o)
        if k:
            s.set_state(pos)   # the same input again (arrange follows integrate or a new state)
        s.arrange()
        compare_grid(s.read_grid(), tra_only if opts else ref)
    s.close()


def test_arrange_reference_default_scene(gsph, po):
    """The reference's own scene (3,944,312 particles, ~296 per cell, SMS)."""
    from pathlib import Path
    path = Path(__file__).resolve().parent.parent / "fast-general-gpu-sph-framework_amd" / \
        "assets" / "scene_default.json"
    p, pos = gsph.scene_json(path)
    banner(f"arrange scene_default n={len(pos)}")
    s = gsph.System(p, pos)
    s.arrange()
    g = s.read_grid()
    import pyoracle
    ref = po.arrange(pyoracle.params_from(p), pos)
    compare_grid(g, ref)
    stats = compare_pair_stats(s, ref)
    print(stats, flush=True)
    T = ref["tasks_count"]
    assert stats["total_pairs"] == (T + 1) // 2 and T > 100000
    s.close()


@pytest.mark.parametrize("split", ["reference", "all_tra"])
def test_reference_scene_block_physics(gsph, po, split):
    """Physics parity at the reference scene's density (assets/
    scene_default.json: h = 0.03, spacing 0.15 h, about 296 particles per
    cell and 1,174 neighbours): a 40^3 block cut from the scene (64,000
    particles, 6^3 dense cells) in the scene's own box, one step against the
    FP64 oracle, under the reference split rule (SMS tasks) and with every
    cell on the TRA path (the split the bench's autotuner picks there: long
    rows keep their accepted-pair masks in the pool).  Both give the same
    bits."""
    from pathlib import Path
    import pyoracle
    path = Path(__file__).resolve().parent.parent / "fast-general-gpu-sph-framework_amd" / \
        "assets" / "scene_default.json"
    p, pos = gsph.scene_json(path)
    keep = np.all((pos[:, :3] >= 0.3) & (pos[:, :3] < 0.48), axis=1)
    blk = np.ascontiguousarray(pos[keep])
    n = len(blk)
    assert 60000 < n < 70000
    op = pyoracle.params_from(p)
    banner(f"reference-scene block {split} n={n}")
    o = gsph.options_default()
    if split == "all_tra":
        for k, v in ALL_TRA.items():
            setattr(o, k, v)
    s = gsph.System(p, blk, options=o)
    s.step(1)
    got = s.read_state()
    info = s.grid_info()
    used, cap = s.mask_pool_info()
    s.close()
    if split == "all_tra":
        assert info["tasks"] == 0 and used > 0      # long TRA rows took the pool path
    else:
        assert info["tasks"] > 100
    ref = po.fp64_step(op, blk, np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32))
    errs = step_errors(*got, ref, op)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= REL_TOL, f"{k} error {v:.3e} > {REL_TOL}"
    # the two splits give the same bits
    o2 = gsph.options_default()
    if split == "reference":
        for k, v in ALL_TRA.items():
            setattr(o2, k, v)
    t = gsph.System(p, blk, options=o2)
    t.step(1)
    other = t.read_state()
    t.close()
    for a, b in zip(got, other):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", SCENES)
def test_one_step_physics(gsph, po, name):
    op, pos = scene(po, name)
    n = len(pos)
    banner(f"step {name} n={n}")
    s = make_system(gsph, op, pos)
    s.step(1)
    pos1, vel1, ev1, acc1 = s.read_state()
    ref = po.fp64_step(op, pos, np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32))
    errs = step_errors(pos1, vel1, ev1, acc1, ref, op)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    if max(errs.values()) > REL_TOL:
        _dump(f"step_{name}", pos=pos, pos1=pos1, vel1=vel1, ev1=ev1, acc1=acc1)
    for k, v in errs.items():
        assert v <= REL_TOL, f"{k} error {v:.3e} > {REL_TOL}"
    s.close()


def test_wall_and_speed_clamps_one_step(gsph, po):
    """Particles that cross a wall in one step (put back at bound -+ (h +
    0.01), velocity reflected by -0.5) and whose speed is clamped to 6 m/s
    (integration_kernels.cu:300-326), next to a resting block: one GPU step
    within the bar of the FP64 oracle, and the clamps' known outcomes."""
    op = make_params(h=0.02, world=(0.64, 0.64, 0.64))
    c = po.derive(op)
    kern = float(np.float32(float(np.float32(0.02)) + 0.01))
    blk = po.scene_block([0.26, 0.26, 0.26], 8, 8, 8, 0.01, 0.01, 11)
    fast, vfast, expect = [], [], []
    k = 0
    for axis in range(3):
        for side in (1, -1):
            bound = float(c.bound_max[axis]) if side > 0 else float(c.bound_min[axis])
            q = [0.1 + 0.07 * k, 0.55 - 0.07 * k, 0.12 + 0.065 * k]
            q[axis] = bound - side * 1e-4
            v = [0.0, 0.0, 0.0]
            v[axis] = 10.0 * side
            fast.append(q + [0.0])
            vfast.append(v)
            expect.append((axis, side, bound))
            k += 1
    # interior particles faster than 6 m/s (speed clamp only)
    fast += [[0.1, 0.5, 0.5, 0.0], [0.5, 0.1, 0.12, 0.0]]
    vfast += [[3.0, -4.0, 5.0], [0.0, 9.0, -2.0]]
    pos = np.concatenate([blk, np.array(fast, np.float32)])
    vel = np.concatenate([np.zeros((len(blk), 3), np.float32), np.array(vfast, np.float32)])
    n = len(pos)
    banner(f"clamps n={n}")
    s = gsph.System(to_gsph_params(gsph, op), pos, vel)
    s.step(1)
    pos1, vel1, ev1, acc1 = s.read_state()
    s.close()
    ref = po.fp64_step(op, pos, vel, np.zeros((n, 4), np.float32))
    errs = step_errors(pos1, vel1, ev1, acc1, ref, op)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for key, v in errs.items():
        assert v <= REL_TOL, f"{key} error {v:.3e} > {REL_TOL}"
    # the sorted slot of each fast particle: its index in the oracle's order
    slot = np.argsort(ref["grid"]["perm"])
    for j, (axis, side, bound) in enumerate(expect):
        i = slot[len(blk) + j]
        assert pos1[i, axis] == np.float32(np.float32(bound) - side * np.float32(kern))
        assert vel1[i, axis] * side < 0
        assert abs(4 * vel1[i, axis] ** 2 + sum(vel1[i, a] ** 2 for a in range(3) if a != axis)
                   - 36.0) < 1e-4
    for j in range(2):
        i = slot[len(blk) + len(expect) + j]
        assert abs(np.linalg.norm(vel1[i].astype(np.float64)) - 6.0) < 1e-5


def _dump(tag, **arrays):
    """Keep the GPU output of a failing numerical case for offline analysis."""
    import os
    from pathlib import Path
    out = Path(os.environ.get("GRAFT_REPO_ROOT", ".")) / "gpurun_out"
    if out.is_dir():
        np.savez_compressed(out / f"{tag}.npz", **arrays)


def test_step_from_moving_state(gsph, po):
    """Physics parity on a state with velocities (viscosity active): run the
    oracle from the GPU's own state after 5 steps."""
    op, pos = mixed_scene(po)
    banner("moving state mixed")
    s = make_system(gsph, op, pos)
    s.step(5)
    p5, v5, e5, _ = s.read_state()
    s.step(1)
    pos1, vel1, ev1, acc1 = s.read_state()
    ref = po.fp64_step(op, p5, v5, e5)
    errs = step_errors(pos1, vel1, ev1, acc1, ref, op)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= REL_TOL, f"{k} error {v:.3e} > {REL_TOL}"
    s.close()


def test_ten_step_checksum_and_grid(gsph, po):
    op, pos = cfg0_params(po)
    banner("10 steps cfg0")
    s = make_system(gsph, op, pos)
    s.step(10)
    pg, vg, eg, _ = s.read_state()
    cs = s.checksum()
    # device FP64 reduction == host sequential FP64 sum of the same state
    np.testing.assert_allclose(cs, po.checksum(pg), rtol=1e-12)
    P, V, E, A = po.step_f32(op, pos, nsteps=10)
    ref = po.checksum(P)
    rel = np.abs(cs - ref) / np.abs(ref)
    print("checksum rel diff", rel, flush=True)
    assert rel.max() <= CHECKSUM_TOL
    # the grid stays bit-exact on the GPU's own trajectory
    s.arrange()
    compare_grid(s.read_grid(), po.arrange(op, pg))
    s.close()


def test_sort_by_merging_matches_oracle(gsph, po):
    """After integrate the slot order is the previous sorted order and only
    the particles whose key changed are sorted again and merged in.  The grid
    must be the oracle's full stable sort of the state, bit for bit: with few
    moved particles, with many (fast random velocities), and right after the
    state is set from the host (every particle treated as moved)."""
    op, pos = mixed_scene(po)
    rng = np.random.default_rng(11)
    for vmax in (0.0, 3.0):
        banner(f"sort by merging vmax={vmax}")
        s = make_system(gsph, op, pos)
        vel = rng.uniform(-vmax, vmax, (len(pos), 3)).astype(np.float32)
        s.set_state(pos, vel)
        s.step(3)
        pg, vg, eg, _ = s.read_state()
        s.arrange()
        g = s.read_grid()
        moved = int(np.count_nonzero(g["perm"] != np.arange(len(pos))))
        print("slots whose particle changed position", moved, flush=True)
        compare_grid(g, po.arrange(op, pg))
        s.density()
        s.force()
        s.integrate()
        # the same state set from the host: a full sort, same grid
        pg2, vg2, eg2, _ = s.read_state()
        s.set_state(pg2, vg2, eg2)
        s.arrange()
        compare_grid(s.read_grid(), po.arrange(op, pg2))
        s.close()


def test_determinism_and_prune_equivalence(gsph, po):
    op, pos = mixed_scene(po)
    banner("determinism mixed")
    out = []
    for prune in (1, 1, 0):
        s = make_system(gsph, op, pos, row_prune=prune)
        s.step(3)
        out.append(s.read_state())
        s.close()
    for k in range(4):
        assert np.array_equal(out[0][k], out[1][k])     # run to run
        assert np.array_equal(out[0][k], out[2][k])     # pruned rows change nothing


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_prune_margins_at_large_coordinates(gsph, po, axis):
    """Row pruning at coordinates near 16000 h (grids up to 16383 cells per
    axis): the FP32 ulp of a coordinate is about 1.5e-3 h there, above the
    fixed prune margins, so the margins scale with the coordinate.  A
    jittered block with cells of 8 particles: pruned and unpruned steps are
    bitwise equal, and the density matches the O(N^2) FP64 brute force."""
    h = 0.02
    world = [24 * h, 24 * h, 24 * h]
    world[axis] = 16380 * h
    op = make_params(h=h, world=tuple(world))
    corner = [4 * h, 4 * h, 4 * h]
    corner[axis] = 16000 * h + 0.0037
    pos = po.scene_block(corner, 16, 16, 16, h / 2, 0.01, 31 + axis)
    banner(f"prune margins axis={axis} n={len(pos)}")
    out = []
    for prune in (1, 0):
        s = make_system(gsph, op, pos, row_prune=prune)
        s.arrange()
        s.density()
        mid = s.read_state()
        s.force()
        s.integrate()
        out.append(mid + s.read_state())
        s.close()
    for k in range(8):
        assert np.array_equal(out[0][k], out[1][k]), k
    rho = po.density_brute_f64(op, out[0][0])
    err = np.abs(out[0][0][:, 3].astype(np.float64) - 1.0 / rho) * rho
    assert err.max() <= REL_TOL, err.max()


ALL_TRA = {"split_mode": 1, "sms_min_self": 1 << 20, "sms_min_total": 1 << 20}


@pytest.mark.parametrize("name,opts", [("mixed", {}), ("dense", ALL_TRA), ("dense", {}),
                                       ("odd_tasks", {}), ("dense", {"sms_mask_bytes": 16})])
def test_pair_masks_change_nothing(gsph, po, name, opts):
    """Force walking density's accepted-pair masks gives the same bits as force
    testing every candidate itself.  Dense all-TRA rows exceed one mask word:
    those tiles record per-round masks in the mask pool, as SMS items do; a
    pool too small for all of them (16 B per particle) makes the rest test
    their candidates in force."""
    op, pos = scene(po, name)
    banner(f"pair masks {name} {opts}")
    out = []
    for masks in (1, 0):
        s = make_system(gsph, op, pos, pair_masks=masks, **opts)
        s.step(3)
        out.append(s.read_state())
        used, cap = s.mask_pool_info()
        if masks and opts == ALL_TRA:
            # dense all-TRA rows exceed one mask word: those tiles keep their
            # masks per staging round in the pool, so the pool path ran
            assert used > 0, (used, cap)
        s.close()
    for k in range(4):
        assert np.array_equal(out[0][k], out[1][k])


@pytest.mark.parametrize("name,opts", [("mixed", {}), ("dense", ALL_TRA)])
@pytest.mark.parametrize("depth,bpc", [(64, 4), (256, 8)])
def test_launch_shape_changes_nothing(gsph, po, name, opts, depth, bpc):
    """Staging depth and persistent grid size leave every bit unchanged."""
    op, pos = scene(po, name)
    banner(f"launch shape {name} depth={depth} bpc={bpc}")
    out = []
    for o in ({}, {"stage_depth": depth, "blocks_per_cu": bpc}):
        s = make_system(gsph, op, pos, **opts, **o)
        s.step(2)
        out.append(s.read_state())
        s.close()
    for k in range(4):
        assert np.array_equal(out[0][k], out[1][k])


@pytest.mark.parametrize("name,opts", [("cfg0", {}), ("mixed", {}), ("dense", ALL_TRA),
                                       ("sparse_random", {}), ("mixed", {"pair_masks": 0})])
def test_lean_passes_change_nothing(gsph, po, name, opts):
    """The lean density/force passes (the TRA tiles without four-chain sums or
    pool rounds, at higher occupancy) and the general passes alone give the
    same bits (the physics tests, which check the FP64 bar, run with the lean
    passes on, the default): cfg0 and sparse_random are taken by the lean
    passes, the dense all-TRA tiles (four chains, pool rounds) and the SMS
    items are left to the general ones, mixed has both, and without pair
    masks the lean force pass leaves every tile."""
    op, pos = scene(po, name)
    banner(f"lean passes {name} {opts}")
    out = []
    for lean in (1, 0):
        s = make_system(gsph, op, pos, lean_passes=lean, **opts)
        s.step(2)
        out.append(s.read_state())
        s.close()
    for k in range(4):
        assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("name", ["mixed", "dense64", "odd_tasks"])
def test_sms_local_merge_changes_no_bits(gsph, po, name):
    """Shared staging of an isSame pair's rows (the reference's SMS local
    merge) on and off: the staged set changes, each lane's pairs and their
    order do not."""
    op, pos = scene(po, name)
    banner(f"sms local merge {name}")
    out = []
    for merge in (1, 0):
        s = make_system(gsph, op, pos, sms_local_merge=merge)
        s.step(3)
        out.append(s.read_state())
        assert s.pair_stats()["same_cell_pairs"] > 0   # the shared staging runs
        s.close()
    for k in range(4):
        assert np.array_equal(out[0][k], out[1][k]), k


def legacy_cut_rows(g, dims):
    """Rows that the legacy SMS staging reads per micro-x layer cut to the
    micro-z bounds (kernel_common.cuh:318-356): rows below (above) the task's
    cell when its z bounds start above 0 (end below 3) and the x-pruned row
    holds more than 22 candidates per layer, counted from the sorted keys."""
    gx, gy, gz = dims
    key = g["hash"].astype(np.int64)
    lc = np.zeros((gx * gy * gz, 4), np.int64)
    np.add.at(lc, (key >> 6, (key >> 4) & 3), 1)
    tasks, T = g["tasks"], g["num_tasks"]
    cut = 0
    for k in range(T):
        t0 = k & ~1
        shared = t0 + 1 < T and tasks[t0]["cellid"] == tasks[t0 + 1]["cellid"]
        tk = tasks[k]
        if shared:
            xi = min(tasks[t0]["xxi"], tasks[t0 + 1]["xxi"])
            xa = max(tasks[t0]["xxx"], tasks[t0 + 1]["xxx"])
            pb = int(tasks[t0]["paired_bounds"])
        else:
            xi, xa, pb = tk["xxi"], tk["xxx"], int(tk["paired_bounds"])
        zmin, zmax = (pb >> 8) & 3, (pb >> 10) & 3
        cx, cy, cz = (int(v) for v in tk["cell_pos"])
        gl, gr = max(4 * cx - 4 + int(xi), 0), min(4 * cx + 4 + int(xa), 4 * gx - 1)
        xs = np.arange(gl, gr + 1)
        for dz in (-1, 1):
            if not ((dz < 0 and zmin > 0) or (dz > 0 and zmax < 3)) or not 0 <= cz + dz < gz:
                continue
            for dy in (-1, 0, 1):
                if 0 <= cy + dy < gy:
                    cnt = lc[(xs >> 2) + gx * (cy + dy + gy * (cz + dz)), xs & 3].sum()
                    cut += cnt > 22 * len(xs)
    return cut


@pytest.mark.parametrize("name", ["mixed", "dense64", "odd_tasks", "dense"])
def test_sms_legacy_staging_changes_no_bits(gsph, po, name):
    """The legacy SMS staging (GSPH_USE_REGISTER_SMS=0: shared rows for every
    isSame pair, dense rows above and below the cell read per layer within
    the task's micro-z bounds) on and off: the candidates it leaves out are
    farther than h from every staged particle and each lane's accepted pairs
    keep their order, so every bit is the same."""
    op, pos = scene(po, name)
    banner(f"sms legacy staging {name}")
    out = []
    for legacy in (1, 0):
        s = make_system(gsph, op, pos, sms_legacy_staging=legacy, sms_local_merge=1 - legacy)
        s.step(3)
        out.append(s.read_state())
        assert s.grid_info()["tasks"] > 0
        s.close()
    for k in range(4):
        assert np.array_equal(out[0][k], out[1][k]), k


def _reference_block(gsph):
    from pathlib import Path
    path = Path(__file__).resolve().parent.parent / "fast-general-gpu-sph-framework_amd" / \
        "assets" / "scene_default.json"
    p, pos = gsph.scene_json(path)
    keep = np.all((pos[:, :3] >= 0.3) & (pos[:, :3] < 0.48), axis=1)
    return p, np.ascontiguousarray(pos[keep])


def test_sms_variants_reference_scene_block(gsph, po):
    """The reference's SMS variants on a 64,000-particle block of the
    reference scene under the reference split (about 296 particles per cell,
    every cell SMS): the legacy staging, which cuts rows there, gives the
    default's bits; the same-cell pair force (every cell has at most 512
    particles, so every SMS particle takes its own cell's pairs from the
    pair pass) stays within REL_TOL of the FP64 oracle, and its pair pass ran
    (the results are not the default's bits)."""
    import pyoracle
    p, blk = _reference_block(gsph)
    n = len(blk)
    op = pyoracle.params_from(p)
    dims = tuple(po.derive(op).grid_size)
    banner(f"sms variants reference-scene block n={n}")
    states = {}
    for tag, opts in (("default", {}), ("legacy", {"sms_legacy_staging": 1}),
                      ("same_cell", {"same_cell_pair_force": 1})):
        o = gsph.options_default()
        for k, v in opts.items():
            setattr(o, k, v)
        s = gsph.System(p, blk, options=o)
        if tag == "legacy":
            s.arrange()
            g = s.read_grid()
            cut = legacy_cut_rows(g, dims)
            print("legacy cut rows", cut, "of", 6 * g["num_tasks"], flush=True)
            assert cut > 100
            s.density()
            s.force()
            s.integrate()
        else:
            s.step(1)
        states[tag] = s.read_state()
        s.close()
    for k in range(4):
        assert np.array_equal(states["legacy"][k], states["default"][k]), k
    assert not np.array_equal(states["same_cell"][3], states["default"][3])
    ref = po.fp64_step(op, blk, np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32))
    errs = step_errors(*states["same_cell"], ref, op)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= REL_TOL, f"{k} error {v:.3e} > {REL_TOL}"


@pytest.mark.parametrize("name", ["dense", "dense64", "odd_tasks", "mixed"])
def test_same_cell_pair_force_physics(gsph, po, name):
    """The same-cell pair force (GSPH_ENABLE_SAME_CELL_PAIR_FORCE): each pair
    inside an SMS cell computed once with both sides accumulated, the SMS
    force skipping its own cell.  A different summation order, so the bar is
    the FP64 oracle at REL_TOL after one step; density is untouched (same
    pressure bits) and the accelerations are not the default's bits (the
    pair pass ran)."""
    op, pos = scene(po, name)
    n = len(pos)
    banner(f"same-cell pair force {name} n={n}")
    s = make_system(gsph, op, pos, same_cell_pair_force=1)
    s.step(1)
    got = s.read_state()
    tasks = s.grid_info()["tasks"]
    s.close()
    assert tasks > 0
    ref = po.fp64_step(op, pos, np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32))
    errs = step_errors(*got, ref, op)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= REL_TOL, f"{k} error {v:.3e} > {REL_TOL}"
    d = make_system(gsph, op, pos)
    d.step(1)
    base = d.read_state()
    d.close()
    assert np.array_equal(got[2][:, 3], base[2][:, 3])          # pressure: density untouched
    assert not np.array_equal(got[3], base[3])                  # the pair pass ran


@pytest.mark.parametrize("name", ["mixed", "dense"])
def test_split_changes_no_bits(gsph, po, name):
    """The SMS/TRA split moves no bit: both paths walk the same nine rows in
    the same (dy, dz) order with candidates in index order, density sums each
    row apart before adding it, and a rejected candidate adds an exact zero
    (an SMS task's wider x-range only adds rejected candidates)."""
    op, pos = scene(po, name)
    banner(f"split neutrality {name}")
    out, tasks = [], []
    for o in ({}, {"split_mode": 1, "sms_min_self": 9, "sms_min_total": 64}, ALL_TRA):
        s = make_system(gsph, op, pos, **o)
        s.step(3)
        out.append(s.read_state())
        tasks.append(s.grid_info()["tasks"])
        s.close()
    assert tasks[0] > 0 and tasks[2] == 0, tasks
    for j in (1, 2):
        for k in range(4):
            assert np.array_equal(out[0][k], out[j][k]), (j, k)


def test_autotune_restores_state_and_keeps_bits(gsph, po):
    op, pos = scene(po, "mixed")
    banner("autotune mixed")
    s = make_system(gsph, op, pos)
    before = s.read_state()
    opt, ms = s.autotune(trial_steps=1)
    after = s.read_state()
    for k in range(3):   # pos, vel, ev (acc is scratch, recomputed every step)
        assert np.array_equal(before[k], after[k])
    assert ms > 0
    assert opt.stage_depth in (64, 128, 256) and opt.blocks_per_cu in (4, 8, 16, 32)
    assert s.options().stage_depth == opt.stage_depth
    s.step(2)
    tuned = s.read_state()
    s.close()
    ref = make_system(gsph, op, pos)
    ref.step(2)
    base = ref.read_state()
    ref.close()
    for k in range(4):
        assert np.array_equal(tuned[k], base[k])


def test_autotuned_split_skewed_scene(gsph, po):
    """Skewed occupancy (pool + 8x denser sphere): the autotuner may move the
    SMS/TRA split; one step stays within tolerance of the FP64 oracle."""
    op, pos = skewed_scene(po)
    n = len(pos)
    banner(f"autotune split skewed n={n}")
    s = make_system(gsph, op, pos)
    opt, ms = s.autotune(trial_steps=1, tune_split=True)
    print("chosen", opt.split_mode, opt.sms_min_self, opt.sms_min_total,
          opt.blocks_per_cu, opt.stage_depth, opt.pair_masks, f"{ms:.3f} ms", flush=True)
    s.step(1)
    pos1, vel1, ev1, acc1 = s.read_state()
    s.close()
    ref = po.fp64_step(op, pos, np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32))
    errs = step_errors(pos1, vel1, ev1, acc1, ref, op)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= REL_TOL, f"{k} error {v:.3e} > {REL_TOL}"


def test_graph_replay_matches_eager(gsph, po):
    op, pos = dense_scene(po)
    banner("graph replay dense")
    a = make_system(gsph, op, pos)
    a.step(4)
    ref = a.read_state()
    b = make_system(gsph, op, pos)
    b.enable_graph(True)
    for _ in range(4):          # replays with host work and syncs in between
        b.step(1)
        b.sync()
    got = b.read_state()
    for k in range(4):
        assert np.array_equal(ref[k], got[k])
    a.close()
    b.close()


def test_stage_timings_and_kernel_times(gsph, po):
    op, pos = cfg0_params(po)
    s = make_system(gsph, op, pos)
    t = s.step(2, timings=True)
    assert len(t) == 2 and all(x["total"] > 0 for x in t)
    s.kernel_timing(True)
    s.step(2)
    kt = s.kernel_times()
    assert kt["force"][1] == 2 and kt["density"][1] == 2 and kt["hash"][1] == 2
    s.kernel_timing(False)
    s.close()


def test_measurement_probes(gsph):
    """The calibration kernels of the roofline figures: the copy probe moves
    its bytes at a plausible HBM rate, the FMA-chain probe reports exactly
    blocks x 256 threads x iterations x 64 lane operations."""
    banner("probes")
    ms = gsph.probe_copy(256 * 2 ** 20, 3)
    gbs = 2 * 256 * 2 ** 20 / (ms * 1e-3) / 1e9
    assert 500 < gbs < 8000, gbs
    ms, ops = gsph.probe_valu(200)
    blocks = ops / (256 * 200 * 64)                      # 8 blocks of 256 per CU
    assert blocks == int(blocks) and int(blocks) % 8 == 0 and blocks >= 8 and ms > 0
    tops = ops / (ms * 1e-3) / 1e12
    assert 5 < tops < 80, tops
    for bad in (0, (1 << 20) + 1):
        with pytest.raises(gsph.GsphError):
            gsph.probe_valu(bad)


def test_tuned_split_same_physics(gsph, po):
    op, pos = mixed_scene(po)
    n = len(pos)
    banner("tuned split mixed")
    s = make_system(gsph, op, pos, split_mode=1, sms_min_self=9, sms_min_total=64)
    s.arrange()
    T_tuned = s.grid_info()["tasks"]
    s.density()
    s.force()
    s.integrate()
    pos1, vel1, ev1, acc1 = s.read_state()
    ref = po.fp64_step(op, pos, np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32))
    assert T_tuned != ref["grid"]["tasks_count"]
    errs = step_errors(pos1, vel1, ev1, acc1, ref, op)
    for k, v in errs.items():
        assert v <= REL_TOL, f"{k} error {v:.3e}"
    s.close()


def test_edge_sizes(gsph, po):
    op = make_params(h=0.02, world=(0.2, 0.2, 0.2))
    banner("edge sizes")
    for pts in ([[0.1, 0.1, 0.1, 0]], [[0.1, 0.1, 0.1, 0], [0.11, 0.1, 0.1, 0]]):
        pos = np.array(pts, np.float32)
        s = make_system(gsph, op, pos)
        s.arrange()
        compare_grid(s.read_grid(), po.arrange(op, pos))
        s.density()
        s.force()
        s.integrate()
        p1, v1, e1, a1 = s.read_state()
        ref = po.fp64_step(op, pos, np.zeros((len(pos), 3), np.float32),
                           np.zeros((len(pos), 4), np.float32))
        assert np.abs(p1[:, 3] * ref["rho"] - 1).max() <= REL_TOL
        s.close()
    # an empty scene steps and reports zeros
    s = make_system(gsph, op, np.zeros((0, 4), np.float32))
    s.step(2)
    s.arrange()
    info = s.grid_info()
    assert info["tasks"] == 0 and info["middle_raw"] == -1
    g = s.read_grid()
    assert g["count"].sum() == 0 and np.all(g["begin"] == 0)
    s.close()


def test_out_of_domain_is_reported(gsph, po):
    op = make_params(h=0.02, world=(0.2, 0.2, 0.2))
    pos = po.scene_block([0.06, 0.06, 0.06], 4, 4, 4, 0.01, 0.0, 1)
    pos[5, 0] = 0.35            # outside the box
    s = make_system(gsph, op, pos)
    s.arrange()
    with pytest.raises(gsph.GsphError) as e:
        s.sync()
    assert e.value.code == -3
    pos[5, 0] = np.nan
    s.set_state(pos)
    s.arrange()
    with pytest.raises(gsph.GsphError) as e:
        s.sync()
    assert e.value.code == -3
    s.close()


def test_api_misuse_and_injected_fault(gsph, po):
    op, pos = dense_scene(po)
    s = make_system(gsph, op, pos)
    with pytest.raises(gsph.GsphError):
        s.density()                     # no grid yet
    s.inject_fault(1)
    s.arrange()                         # call 1 passes
    with pytest.raises(gsph.GsphError) as e:
        s.density()                     # call 2 fails, launches nothing
    assert e.value.code == -9
    s.density()
    s.close()


def test_batch_alone_vs_mix(gsph, po):
    """Scenes batched into one context give bitwise the results they give alone."""
    rng = np.random.default_rng(21)
    op = make_params(h=0.02, world=(1, 1, 1))
    worlds, chunks = [], []
    for k in range(5):
        nx, ny, nz = rng.integers(4, 14, 3)
        pos = po.scene_block([0.06, 0.06, 0.06], int(nx), int(ny), int(nz), 0.01, 0.01, 100 + k)
        if k == 2:   # a dense scene (SMS tasks)
            pos = po.scene_block([0.06, 0.06, 0.06], 12, 12, 12, 0.005, 0.01, 100 + k)
        ext = pos[:, :3].max(0) - 0.06 + 0.01
        worlds.append([max(0.2, 2 * e + 0.12) for e in ext])
        chunks.append(pos)
    counts = [len(c) for c in chunks]
    gp = to_gsph_params(gsph, op)
    banner(f"batch alone vs mix counts={counts}")
    b = gsph.System(gp, np.concatenate(chunks), batch_worlds=worlds, batch_counts=counts)
    b.step(3)
    bp, bv, be, ba = b.read_state()
    for k, c in enumerate(chunks):
        gp1 = to_gsph_params(gsph, op)
        for a in range(3):
            gp1.world_size[a] = worlds[k][a]
        s = gsph.System(gp1, c)
        s.step(3)
        sp, sv, se, sa = s.read_state()
        f, cnt = b.scene_range(k)
        assert cnt == len(c)
        for x, y in ((bp, sp), (bv, sv), (be, se), (ba, sa)):
            assert np.array_equal(x[f:f + cnt], y)
        cs_b = b.checksum(f, cnt)
        assert np.array_equal(cs_b, s.checksum())
        s.close()
    b.close()


def test_batch_tiny_scenes_alone_vs_mix(gsph, po):
    """Scenes smaller than one 64-particle tile (several scenes start inside
    one tile, and inside one 64-particle block of the layout's scene table)
    give bitwise the results they give alone."""
    op = make_params(h=0.02, world=(1, 1, 1))
    dims = [(2, 2, 2), (1, 1, 3), (3, 3, 3), (1, 1, 1), (4, 4, 4), (2, 3, 4), (5, 5, 5), (1, 2, 1)]
    worlds, chunks = [], []
    for k, (nx, ny, nz) in enumerate(dims):
        pos = po.scene_block([0.06, 0.06, 0.06], nx, ny, nz, 0.01, 0.01, 300 + k)
        ext = pos[:, :3].max(0) - 0.06 + 0.01
        worlds.append([max(0.2, 2 * e + 0.12) for e in ext])
        chunks.append(pos)
    counts = [len(c) for c in chunks]
    gp = to_gsph_params(gsph, op)
    banner(f"batch tiny scenes counts={counts}")
    b = gsph.System(gp, np.concatenate(chunks), batch_worlds=worlds, batch_counts=counts)
    b.step(3)
    bp, bv, be, ba = b.read_state()
    for k, c in enumerate(chunks):
        gp1 = to_gsph_params(gsph, op)
        for a in range(3):
            gp1.world_size[a] = worlds[k][a]
        s = gsph.System(gp1, c)
        s.step(3)
        sp, sv, se, sa = s.read_state()
        f, cnt = b.scene_range(k)
        assert cnt == len(c)
        for x, y in ((bp, sp), (bv, sv), (be, se), (ba, sa)):
            assert np.array_equal(x[f:f + cnt], y)
        s.close()
    b.close()


def test_config1_full_size_properties(gsph, po):
    """Workload 1 at full size (3,944,312 particles, 317^3 cells): the sort
    against the oracle's keys, the first step against the FP64 oracle, the
    checksum after 10 steps against the FP32 CPU step, and the sort by
    merging on later steps."""
    p, pos = gsph.scene_config(1)
    n = len(pos)
    banner(f"config1 n={n}")
    s = gsph.System(p, pos)
    s.arrange()
    g = s.read_grid()
    perm = g["perm"]
    assert np.array_equal(np.sort(perm), np.arange(n))
    import pyoracle
    key, bad = po.hash_(pyoracle.params_from(p), pos)
    assert bad == 0
    assert np.array_equal(g["hash"], key[perm])
    assert np.array_equal(perm, np.lexsort((np.arange(n), key)))
    cnt = g["count"]
    assert cnt.sum() == n
    assert np.array_equal(g["begin"][1:], np.cumsum(cnt)[:-1]) and g["begin"][0] == 0
    assert np.array_equal(g["compact"], np.arange(n))      # all TRA at 8 per cell
    assert g["tasks"].shape[0] == 0 and g["middle_raw"] == -1
    s.density()
    s.force()
    s.integrate()
    pos1, vel1, ev1, acc1 = s.read_state()
    ref = po.fp64_step(pyoracle.params_from(p), pos, np.zeros((n, 3), np.float32),
                       np.zeros((n, 4), np.float32))
    errs = step_errors(pos1, vel1, ev1, acc1, ref, pyoracle.params_from(p))
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= REL_TOL, f"{k} error {v:.3e}"
    del ref
    # ten steps: the checksum against the FP32 CPU step of the reference
    s.step(9)
    cs = s.checksum()
    P, _, _, _ = po.step_f32(pyoracle.params_from(p), pos, nsteps=10)
    rel = np.abs(cs - po.checksum(P)) / np.abs(po.checksum(P))
    print("checksum rel diff", rel, flush=True)
    assert rel.max() <= CHECKSUM_TOL
    # later steps sort by merging the moved particles into the previous
    # order: still the stable sort of the keys of the state in slot order
    s.step(3)
    pg, _, _, _ = s.read_state()
    key, bad = po.hash_(pyoracle.params_from(p), pg)
    assert bad == 0
    s.arrange()
    g = s.read_grid()
    assert np.array_equal(g["perm"], np.lexsort((np.arange(n), key)))
    assert np.array_equal(g["hash"], key[g["perm"]])
    s.close()


def test_config2_full_size(gsph, po):
    """Workload 2 at full size (8,137,432 particles: a 0.01-spaced pool and a
    0.005-spaced sphere, 0 to 64 per cell): arrange bit-exact against the
    oracle under the reference split rule, one step within tolerance of the
    FP64 oracle, and the autotuned split giving the same bits."""
    import pyoracle
    p, pos = gsph.scene_config(2)
    n = len(pos)
    banner(f"config2 n={n}")
    op = pyoracle.params_from(p)
    s = gsph.System(p, pos)
    s.arrange()
    g = s.read_grid()
    ref = po.arrange(op, pos)
    compare_grid(g, ref)
    st = compare_pair_stats(s, ref)
    print(st, flush=True)
    assert ref["tasks_count"] > 100000 and 0 < ref["middle_raw"] < n
    s.density()
    s.force()
    s.integrate()
    pos1, vel1, ev1, acc1 = s.read_state()
    s.close()
    ref = po.fp64_step(op, pos, np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32))
    errs = step_errors(pos1, vel1, ev1, acc1, ref, op)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= REL_TOL, f"{k} error {v:.3e}"
    t = gsph.System(p, pos)
    opt, ms = t.autotune(trial_steps=1, tune_split=True)
    print("chosen", opt.split_mode, opt.sms_min_self, opt.sms_min_total, opt.blocks_per_cu,
          opt.stage_depth, opt.pair_masks, f"{ms:.3f} ms", flush=True)
    t.step(1)
    got = t.read_state()
    t.close()
    for k, a in enumerate((pos1, vel1, ev1, acc1)):
        assert np.array_equal(got[k], a), k


def test_config3_full_size_batch(gsph, po):
    """Workload 3 at full size (256 scenes, 17,280,876 particles, autotuned,
    hipGraph step): sampled scenes give bitwise what they give alone, and the
    first step of two of them is within tolerance of the FP64 oracle."""
    import pyoracle
    worlds, blocks = gsph.batch_workload(256)
    counts = [len(b) for b in blocks]
    banner(f"config3 scenes=256 n={sum(counts)}")
    params, _ = gsph.scene_config(0)
    b = gsph.System(params, np.concatenate(blocks), batch_worlds=worlds, batch_counts=counts)
    b.autotune(trial_steps=1)
    b.enable_graph(True)
    b.step(1)
    first = b.read_state()
    b.step(2)
    state = b.read_state()
    order = np.argsort(counts)
    rng = np.random.default_rng(3)
    sample = [int(order[0]), int(order[-1])] + [int(k) for k in rng.choice(256, 4, replace=False)]
    for k in sample:
        p, _ = gsph.scene_config(0)
        for a in range(3):
            p.world_size[a] = worlds[k][a]
        s = gsph.System(p, blocks[k])
        s.step(3)
        alone = s.read_state()
        f, cnt = b.scene_range(k)
        assert cnt == counts[k]
        for x, y in zip(state, alone):
            assert np.array_equal(x[f:f + cnt], y), k
        assert np.array_equal(b.checksum(f, cnt), s.checksum())
        s.close()
    for k in sample[:2]:
        p, _ = gsph.scene_config(0)
        for a in range(3):
            p.world_size[a] = worlds[k][a]
        op = pyoracle.params_from(p)
        f, cnt = b.scene_range(k)
        ref = po.fp64_step(op, blocks[k], np.zeros((cnt, 3), np.float32),
                           np.zeros((cnt, 4), np.float32))
        errs = step_errors(*(x[f:f + cnt] for x in first), ref, op)
        print(k, cnt, {n: f"{v:.2e}" for n, v in errs.items()}, flush=True)
        for n, v in errs.items():
            assert v <= REL_TOL, f"scene {k} {n} error {v:.3e}"
    b.close()
