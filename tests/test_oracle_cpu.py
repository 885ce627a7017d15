"""CPU tests that pin the oracle (oracle/oracle.cpp) without a GPU.

Each check is chosen so that a plausible restatement mistake fails it:
known-answer cases computed by hand, closed forms, invariants, brute force on
tiny inputs, and an independent numpy restatement of the bit-exact integer
rules.  The reference ships no tests or golden vectors, so
these are the pins.
"""
import math

import numpy as np
import pytest

from helpers import cfg0_params, make_params


# ---------------------------------------------------------------- scene inputs
def test_splitmix64_published_vector(po):
    # first outputs of SplitMix64 for seed 1234567 (the public reference demo)
    assert po.splitmix64(1234567, 5) == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]


def test_scene_block_layout(po):
    pos = po.scene_block([0.5, 0.25, 0.125], 2, 3, 4, 0.01, 0.0, 1)
    assert pos.shape == (24, 4)
    # x outer, z inner; (i + 1/2) * spacing from the corner; w = 0
    assert pos[0, :3].tolist() == pytest.approx([0.505, 0.255, 0.13], abs=1e-7)
    assert pos[1, :3].tolist() == pytest.approx([0.505, 0.255, 0.14], abs=1e-7)
    assert pos[4, :3].tolist() == pytest.approx([0.505, 0.265, 0.13], abs=1e-7)
    assert pos[12, :3].tolist() == pytest.approx([0.515, 0.255, 0.13], abs=1e-7)
    assert np.all(pos[:, 3] == 0)
    jit = po.scene_block([0.5, 0.25, 0.125], 2, 3, 4, 0.01, 0.01, 1)
    d = jit[:, :3] - pos[:, :3]
    assert np.all(np.abs(d) <= 0.01 * 0.01 * (1 + 1e-3))
    assert np.abs(d).max() > 0


def test_reference_default_scene_count(po):
    # assets/scene_default.json: block 0.091..0.8, h = 0.03, interval 0.15;
    # the reference reports 3,944,312 = 158^3 particles for it.
    pos = po.scene_lattice([0.091] * 3, [0.8] * 3, 0.03, 0.15)
    assert pos.shape[0] == 158 ** 3 == 3944312
    # the FP32 accumulation x += 0.0045f is reproduced, not recomputed
    x = np.float32(0.091)
    xs = []
    while x < np.float32(0.8):
        xs.append(x)
        x = np.float32(x + np.float32(np.float32(0.03) * np.float32(0.15)))
    assert np.array_equal(np.unique(pos[:, 0]), np.array(sorted(set(xs)), np.float32))


# ---------------------------------------------------------------- constants
def test_derived_constants_closed_form(po):
    p = make_params(h=0.02, mass=1e-3, world=(0.64, 0.64, 0.64))
    c = po.derive(p)
    h = np.float64(np.float32(0.02))
    assert c.grid_size[:] == [32, 32, 32]
    assert c.inv_cell_size == np.float32(1) / np.float32(0.02)
    assert c.poly6_value == np.float32(315 / (64 * math.pi * h ** 9))
    assert c.spiky_value == np.float32(-45 / (math.pi * h ** 6))
    assert c.half_spiky_value == np.float32(c.spiky_value) * np.float32(0.5)
    assert c.density_scale == np.float32(np.float32(1e-3) * np.float32(c.poly6_value))
    # self density m * poly6 * h^6 = m * 315 / (64 pi h^3)
    assert c.self_density == pytest.approx(1e-3 * 315 / (64 * math.pi * h ** 3), rel=1e-6)
    assert c.bound_min[:] == [np.float32(0.02)] * 3
    assert c.bound_max[0] == np.float32(np.float32(0.64) - np.float32(0.02))


# ---------------------------------------------------------------- micro hash
def _numpy_hash(pos, h, grid):
    """Independent numpy restatement of sph_utils.cuh:53-105 (IEEE float32)."""
    rat = np.float32(4.0) * (np.float32(1.0) / np.float32(h))
    m = np.floor(pos[:, :3].astype(np.float32) * rat).astype(np.int64)
    gx, gy, gz = grid
    ok = (m >= 0).all(1) & (m[:, 0] < 4 * gx) & (m[:, 1] < 4 * gy) & (m[:, 2] < 4 * gz)
    c = (m[:, 0] >> 2) + gx * ((m[:, 1] >> 2) + gy * (m[:, 2] >> 2))
    idi = (m[:, 1] & 3) + (((m[:, 2] & 3) + ((m[:, 0] & 3) << 2)) << 2)
    return np.where(ok, (c << 6) | idi, -1).astype(np.int32)


def test_hash_known_answers(po):
    p = make_params(h=0.02, world=(0.64, 0.64, 0.64))
    pos = np.array([[0.0651, 0.0123, 0.1987, 0],    # micro (13, 2, 39)
                    [0.025, 0.025, 0.025, 0],       # 0.025f * 200 rounds to 5.0
                    [0.6399, 0.0, 0.0, 0],          # last micro column in x
                    [0.64, 0.1, 0.1, 0],            # outside the grid -> -1
                    [-0.001, 0.1, 0.1, 0]], np.float32)
    key, bad = po.hash_(p, pos)
    # (13,2,39): coarse (3,0,9) -> 3 + 32*(0 + 32*9) = 9219; idi = 2 + 4*3 + 16*1 = 30
    assert key[0] == 9219 * 64 + 30
    # (5,5,5): coarse (1,1,1) = 1 + 32 + 1024 = 1057; idi = 1 + 4 + 16 = 21
    assert key[1] == 1057 * 64 + 21
    # (127, 0, 0): coarse (31,0,0); idi = 16*3 = 48
    assert key[2] == 31 * 64 + 48
    assert key[3] == -1 and key[4] == -1 and bad == 2


def test_hash_matches_numpy_restatement(po):
    rng = np.random.default_rng(5)
    p = make_params(h=0.02, world=(0.64, 0.40, 0.52))
    pos = np.zeros((20000, 4), np.float32)
    pos[:, :3] = rng.uniform(0, 1, (20000, 3)) * np.array([0.64, 0.40, 0.52])
    # include exact micro boundaries and their float neighbours
    edge = (np.arange(1, 120) * np.float32(0.005)).astype(np.float32)
    pos[:119, 0] = edge
    pos[119:238, 0] = np.nextafter(edge, np.float32(0))
    key, _ = po.hash_(p, pos)
    assert np.array_equal(key, _numpy_hash(pos, 0.02, (32, 20, 26)))


# ---------------------------------------------------------------- split rule
@pytest.mark.parametrize("n,neigh,expect", [
    (14, 0, 0),      # n < 15 and total < 96 -> TRA
    (15, 0, 1),      # n = 15: total < 60 -> (15 + 27) >> 5 = 1
    (32, 0, 1),      # (32 + 27) >> 5 = 1
    (33, 0, 1),      # (33 + 27) >> 5 = 1   (total 33 < 60)
    (36, 0, 1),      # (36 + 27) >> 5 = 1
    (37, 0, 2),      # (37 + 27) >> 5 = 2
    (33, 30, 2),     # total 63 >= 60 -> (33 + 31) >> 5 = 2
    (14, 82, 1),     # total 96: not TRA any more -> (14 + 31) >> 5 = 1
    (14, 81, 0),     # total 95 < 96 and n < 15 -> TRA
    (1, 95, 1),      # one particle, total 96 -> a task of one particle
    (0, 500, 0),     # empty cell
])
def test_split_rule_known_answers(po, n, neigh, expect):
    grid = (3, 3, 3)
    count = np.zeros(27, np.int32)
    centre = 1 + 3 * (1 + 3 * 1)
    count[centre] = n
    # spread the neighbour total over the 6 face neighbours
    faces = [centre - 1, centre + 1, centre - 3, centre + 3, centre - 9, centre + 9]
    for k in range(neigh):
        count[faces[k % 6]] += 1
    # a corner (non-face) neighbour must not count
    count[0] += 100
    req = po.split_req(count, grid)
    assert req[centre] == expect


def test_split_rule_boundary_cells(po):
    # face neighbours outside the grid are skipped, not wrapped
    grid = (4, 1, 1)
    count = np.array([20, 0, 0, 90], np.int32)
    req = po.split_req(count, grid)
    assert req.tolist() == [1, 0, 0, 3]


# ---------------------------------------------------------------- arrange
def _check_arrange_invariants(po, p, pos, g):
    n = pos.shape[0]
    c = po.derive(p)
    key, _ = po.hash_(p, pos)
    perm = g["perm"]
    assert np.array_equal(np.sort(perm), np.arange(n))              # bijection
    assert np.array_equal(g["hash"], key[perm])                      # sorted keys
    order = np.lexsort((np.arange(n), key))                          # stable order
    assert np.array_equal(perm, order)
    assert g["count"].sum() == n
    ex = np.concatenate([[0], np.cumsum(g["count"])[:-1]])
    assert np.array_equal(g["begin"], ex)
    cell = g["hash"] >> 6
    assert np.array_equal(np.bincount(cell, minlength=len(g["count"])), g["count"])
    grid = c.grid_size[:]
    assert np.array_equal(g["req"], po.split_req(g["count"], grid))
    # TRA/SMS partition
    rank = np.arange(n) - g["begin"][cell]
    sms = rank < 32 * g["req"][cell]
    tra_idx = np.nonzero(~sms)[0]
    sms_idx = np.nonzero(sms)[0]
    assert np.array_equal(g["compact"], np.concatenate([tra_idx, sms_idx]))
    assert g["middle_raw"] == (len(tra_idx) if len(sms_idx) else -1)
    assert g["tasks_count"] == g["req"].sum()


def test_arrange_invariants_random(po):
    rng = np.random.default_rng(11)
    p = make_params(h=0.02, world=(0.3, 0.2, 0.25))
    # clustered cloud: dense blobs (SMS) plus sparse background (TRA)
    pts = [rng.uniform([0.02] * 3, [0.28, 0.18, 0.23], (3000, 3))]
    for centre in ([0.1, 0.1, 0.1], [0.2, 0.05, 0.15]):
        pts.append(rng.normal(centre, 0.01, (4000, 3)))
    xyz = np.clip(np.concatenate(pts), [0.021] * 3, [0.279, 0.179, 0.229])
    pos = np.zeros((len(xyz), 4), np.float32)
    pos[:, :3] = xyz
    g = po.arrange(p, pos)
    _check_arrange_invariants(po, p, pos, g)
    assert g["tasks_count"] > 10 and g["middle_raw"] > 0


def test_arrange_tasks_known_answer(po):
    """One cell with 70 particles spread over micro layers -> 3 tasks.

    Cell (1,1,1) of a 3^3 grid, h = 0.04 (micro cell 0.01).  Particles are
    placed so that sorted order is known: 20 particles in micro (x=0,z=0,y=0),
    20 in (x=0,z=1,y=2), 30 in (x=2,z=3,y=3).
    """
    p = make_params(h=0.04, world=(0.12, 0.12, 0.12))
    base = 0.04
    groups = [((0, 0, 0), 20), ((0, 2, 1), 20), ((2, 3, 3), 30)]  # (mx, my, mz)
    rows = []
    for (mx, my, mz), k in groups:
        for j in range(k):
            rows.append([base + 0.01 * mx + 0.0005 + 1e-5 * j,
                         base + 0.01 * my + 0.005, base + 0.01 * mz + 0.005, 0])
    pos = np.array(rows, np.float32)[::-1].copy()   # reversed input order
    g = po.arrange(p, pos)
    _check_arrange_invariants(po, p, pos, g)
    centre = 1 + 3 * (1 + 3 * 1)
    assert g["count"][centre] == 70 and g["req"][centre] == 3   # (70 + 27) >> 5
    t = g["tasks"]
    assert g["tasks_count"] == 3 and len(t) == 4                # odd -> pad
    # task 0: particles 0..31: first in (0,0,0), last in (0,y2,z1): same x
    assert (t[0]["xxi"], t[0]["xxx"], t[0]["zzi"], t[0]["zzz"]) == (0, 0, 0, 1)
    assert (t[0]["yyi"], t[0]["yyy"]) == (0, 3)                 # z differs -> full y
    # task 1: particles 32..63: (0,2,1) .. (2,3,3): x differs -> full y, z
    assert (t[1]["xxi"], t[1]["xxx"], t[1]["yyi"], t[1]["yyy"], t[1]["zzi"], t[1]["zzz"]) == (0, 2, 0, 3, 0, 3)
    # task 2: particles 64..69, all in (2,3,3): x, z and y bounds exact
    assert (t[2]["xxi"], t[2]["xxx"], t[2]["yyi"], t[2]["yyy"], t[2]["zzi"], t[2]["zzz"]) == (2, 2, 3, 3, 3, 3)
    # judgeTask: pair (0,1) same cell -> isSame; first.xxi (0) != second.xxx (2)
    # -> widened: z bounds (0,3); bounds pack (x 0..2, y from first, z 0..3)
    assert t[0]["is_same"] == 1 and t[1]["is_same"] == 1
    pb = int(t[0]["paired_bounds"])
    assert (pb & 3, (pb >> 2) & 3, (pb >> 4) & 3, (pb >> 6) & 3, (pb >> 8) & 3, (pb >> 10) & 3) == (0, 2, 0, 3, 0, 3)
    # odd tail: task 2 isSame, pad copies the cell with p_offset + 32
    assert t[2]["is_same"] == 1 and t[3]["is_same"] == 1
    assert t[3]["cellid"] == centre and t[3]["p_offset"] == 96
    assert t[3]["cell_begin"] == t[2]["cell_begin"] and t[3]["cell_particle_count"] == 70
    for k in range(3):
        assert t[k]["p_offset"] == 32 * k and t[k]["cellid"] == centre
        assert list(t[k]["cell_pos"]) == [1, 1, 1]


def test_judge_compact_pair(po):
    """Two tasks whose x bounds meet (first.xxi == second.xxx) keep z bounds."""
    p = make_params(h=0.04, world=(0.12, 0.12, 0.12))
    rows = []
    for j in range(64):   # all in micro x = 1; z from 0 to 3 along j
        mz = j // 16
        rows.append([0.04 + 0.015, 0.04 + 0.005 + 0.01 * (j % 4), 0.04 + 0.01 * mz + 0.005, 0])
    pos = np.array(rows, np.float32)
    g = po.arrange(p, pos)
    t = g["tasks"]
    assert g["tasks_count"] == 2 and t[0]["is_same"] == 1
    pb = int(t[0]["paired_bounds"])
    # first z bounds 0..1 (particles 0..31), second 2..3 -> compact pair z 0..3
    assert (pb >> 8) & 3 == t[0]["zzi"] and (pb >> 10) & 3 == t[1]["zzz"]
    stats = dict(compact=int(t[0]["xxi"] == t[1]["xxx"]))
    assert stats["compact"] == 1


def test_arrange_lattice_all_tra(po):
    p, pos = cfg0_params(po)
    g = po.arrange(p, pos)
    _check_arrange_invariants(po, p, pos, g)
    assert g["tasks_count"] == 0 and g["middle_raw"] == -1
    assert np.array_equal(g["compact"], np.arange(len(pos)))


# ---------------------------------------------------------------- physics
def test_density_closed_forms(po):
    p = make_params(h=0.02, mass=1e-3, world=(0.2, 0.2, 0.2))
    h = float(np.float32(0.02))
    m = float(np.float32(1e-3))
    poly6 = 315 / (64 * math.pi * h ** 9)
    r = 0.013
    pos = np.array([[0.1, 0.1, 0.1, 0], [0.1 + r, 0.1, 0.1, 0], [0.05, 0.15, 0.05, 0]], np.float32)
    rho, pr = po.density_f64(p, pos)
    r32 = float(np.float32(0.1 + r)) - float(np.float32(0.1))
    pair = m * poly6 * ((h * h - r32 * r32) ** 3 + h ** 6)
    assert rho[0] == pytest.approx(pair, rel=1e-12)
    assert rho[1] == pytest.approx(pair, rel=1e-12)
    assert rho[2] == pytest.approx(m * poly6 * h ** 6, rel=1e-12)    # isolated
    assert pr[2] == pytest.approx((rho[2] / 1000.0) ** 7 - 1.0, rel=1e-12)


def test_density_grid_vs_brute_force(po):
    rng = np.random.default_rng(3)
    p = make_params(h=0.02, world=(0.16, 0.16, 0.16))
    pos = np.zeros((1500, 4), np.float32)
    pos[:, :3] = rng.uniform(0.021, 0.139, (1500, 3))
    rho, _ = po.density_f64(p, pos)
    rho_b = po.density_brute_f64(p, pos)
    np.testing.assert_allclose(rho, rho_b, rtol=1e-12)


def test_force_two_particle_closed_form(po):
    p = make_params(h=0.02, mass=1e-3, world=(0.2, 0.2, 0.2))
    h = float(np.float32(0.02))
    m = float(np.float32(1e-3))
    pos = np.array([[0.1, 0.1, 0.1, 0], [0.11, 0.1, 0.1, 0]], np.float32)
    ev = np.array([[0.0, 0.5, 0.0, 0], [0.0, -0.5, 0.0, 0]], np.float32)
    rho = np.array([900.0, 1100.0])
    pr = np.array([0.3, -0.2])
    acc, _ = po.force_f64(p, pos, ev, rho, pr)
    r = float(np.float32(0.11)) - float(np.float32(0.1))
    kr = h - r
    half_spiky = -45 / (math.pi * h ** 6) / 2
    visc = 6.5 * 45 / (math.pi * h ** 6)
    # particle 0: rel = x0 - x1 = (-r, 0, 0); V = 1/rho1
    V = 1 / rho[1]
    P = -(-r) * V * (pr[0] + pr[1]) * kr * kr / r
    Vis = (-0.5 - 0.5) * V * kr
    t = h * h - r * r
    G = -r * V * t * t * (-945 / (32 * math.pi * h ** 9)) * m
    L = V * t * (r * r - 0.75 * t) * (945 / (8 * math.pi * h ** 9)) * m / rho[0]
    coe = float(np.float32(0.2))
    surf_x = G * coe * L / abs(G) if abs(G) > float(np.float32(0.1)) else 0.0
    assert acc[0, 0] == pytest.approx(m * P * half_spiky + surf_x, rel=1e-10)
    assert acc[0, 1] == pytest.approx(m * Vis * visc, rel=1e-10)
    assert acc[0, 2] == 0.0


@pytest.mark.parametrize("delta,near", [(1e-7, True), (-1e-7, True), (1e-3, False), (-1e-3, False)])
def test_force_branch_alternative_on_surface_threshold(po, delta, near):
    """Two particles, rho_1 chosen so that particle 0's |grad c| is
    (1 + delta) x surface_normal: within FP32 reach of the threshold the
    oracle also returns the other branch (surface force toggled), elsewhere
    the same acceleration."""
    p = make_params(h=0.02, mass=1e-3, world=(0.2, 0.2, 0.2))
    h = float(np.float32(0.02))
    m = float(np.float32(1e-3))
    pos = np.array([[0.1, 0.1, 0.1, 0], [0.11, 0.1, 0.1, 0]], np.float32)
    r = float(np.float32(0.11)) - float(np.float32(0.1))
    t = h * h - r * r
    gscale = 945 / (32 * math.pi * h ** 9) * m
    thr = float(np.float32(0.1))
    V = thr * (1 + delta) / (r * t * t * gscale)
    rho = np.array([900.0, 1 / V])
    pr = np.array([0.3, -0.2])
    acc, scale, alt = po.force_f64(p, pos, np.zeros_like(pos), rho, pr, alt=True)
    kr = h - r
    base = m * (r * V * (pr[0] + pr[1]) * kr * kr / r) * (-45 / (math.pi * h ** 6) / 2)
    G = r * V * t * t * gscale            # rel = -r and grad_poly6 < 0: G_x > 0
    L = V * t * (r * r - 0.75 * t) * (945 / (8 * math.pi * h ** 9)) * m / rho[0]
    surf = float(np.float32(0.2)) * L     # G coe L / |G| along +x
    on = delta > 0
    assert acc[0, 0] == pytest.approx(base + (surf if on else 0.0), rel=1e-9)
    other = base + (0.0 if on else surf)
    assert alt[0, 0] == pytest.approx(other if near else acc[0, 0], rel=1e-9)
    assert abs(surf) > 1e-3 * abs(base)   # the two branches differ visibly
    assert np.array_equal(alt[1], acc[1])  # |grad c| of particle 1 is 0.18 x threshold
    assert np.array_equal(alt[0, 1:], acc[0, 1:])


def test_force_symmetric_lattice_interior_is_balanced(po):
    # dyadic spacing and corner: every lattice coordinate is exact in float32
    s = 2.0 ** -7
    p = make_params(h=2 * s, world=(0.5, 0.5, 0.5))
    pos = po.scene_block([0.125, 0.125, 0.125], 12, 12, 12, s, 0.0, 1)
    rho, pr = po.density_f64(p, pos)
    acc, scale = po.force_f64(p, pos, np.zeros_like(pos), rho, pr)
    centre = np.all(np.abs(pos[:, :3] - (0.125 + 6 * s)) < 1.2 * s, axis=1)
    assert centre.sum() == 8
    assert np.abs(acc[centre]).max() < 1e-9 * scale[centre].max() + 1e-12


def test_integrate_free_fall_and_walls(po):
    p = make_params(h=0.02, world=(1.0, 1.0, 1.0), dt=1e-3)
    pos = np.array([[0.5, 0.5, 0.5, 0.001], [0.5, 0.021, 0.5, 0.001]], np.float32)
    vel = np.array([[0.1, 0.0, 0.0], [0.0, -1.0, 0.0]], np.float32)
    ev = np.array([[0.2, 0.0, 0.0, 7.0], [0.0, 0.0, 0.0, 0.0]], np.float32)
    acc = np.zeros((2, 3))
    rho = np.array([1000.0, 1000.0])
    po_, vo, eo = po.integrate_f64(p, pos, vel, ev, acc, rho)
    g = float(np.float32(-9.81))
    dt = float(np.float32(1e-3))
    assert vo[0].tolist() == pytest.approx([float(np.float32(0.1)), g * dt, 0.0], rel=1e-12)
    assert po_[0, 1] == pytest.approx(float(np.float32(0.5)) + g * dt * dt, rel=1e-12)
    assert eo[0, 0] == pytest.approx((float(np.float32(0.1)) + float(np.float32(0.2))) / 2)
    assert eo[0, 3] == 7.0
    # particle 1 sits in the floor penalty band: a_y = (alpha diff - beta v) / rho + g,
    # clamped to |a| <= sqrt(3e6); it moves up
    kern = float(np.float32(float(np.float32(0.02)) + 0.01))
    diff = kern - (float(np.float32(0.021)) - float(np.float32(0.02)))
    ay = (2e7 * diff - 2e5 * -1.0) / 1000.0 + g
    ay = min(ay, math.sqrt(3e6))
    assert vo[1, 1] == pytest.approx(-1.0 + ay * dt, rel=1e-9)


def test_integrate_speed_clamp_closed_form(po):
    """|v|^2 > 36 after the Euler update: the speed is scaled to exactly 6
    with the direction kept (integration_kernels.cu:300-306).  Far from every
    wall, no force: v' = v + g dt, then v' * 6 / |v'|."""
    p = make_params(h=0.02, world=(1.0, 1.0, 1.0), dt=1e-4)
    pos = np.array([[0.5, 0.5, 0.5, 0.001], [0.4, 0.6, 0.5, 0.001]], np.float32)
    vel = np.array([[0.0, 0.0, 10.0], [3.0, -4.0, 5.0]], np.float32)   # |v|^2 = 100, 50
    ev = np.zeros((2, 4), np.float32)
    po_, vo, eo = po.integrate_f64(p, pos, vel, ev, np.zeros((2, 3)), np.array([1000.0, 1000.0]))
    g = float(np.float32(-9.81))
    dt = float(np.float32(1e-4))
    for i in range(2):
        v1 = vel[i].astype(np.float64) + np.array([0.0, g * dt, 0.0])
        assert np.linalg.norm(vo[i]) == pytest.approx(6.0, rel=1e-14)
        np.testing.assert_allclose(vo[i] / 6.0, v1 / np.linalg.norm(v1), rtol=1e-14, atol=1e-16)
        np.testing.assert_allclose(po_[i, :3], pos[i, :3].astype(np.float64) + vo[i] * dt,
                                   rtol=1e-15)
        np.testing.assert_allclose(eo[i, :3], vo[i] / 2, rtol=1e-15)


@pytest.mark.parametrize("axis,side", [(0, 1), (1, -1), (2, 1), (2, -1)])
def test_integrate_wall_clamp_closed_form(po, axis, side):
    """A particle that crosses a wall in one step is put back at
    bound -+ (h + 0.01) and its velocity component reflected by
    wall_damping = -0.5 (integration_kernels.cu:308-326).  Heading into the
    wall at 10 m/s from 1e-4 m inside it, the speed clamp leaves |v| = 6
    before the wall, so 4 v_axis'^2 + (other components)^2 = 36 after it."""
    h = 0.02
    p = make_params(h=h, world=(1.0, 1.0, 1.0), dt=1e-4)
    c = po.derive(p)
    kern = float(np.float32(float(np.float32(h)) + 0.01))
    bound = float(c.bound_max[axis]) if side > 0 else float(c.bound_min[axis])
    pos = np.array([[0.5, 0.5, 0.5, 0.0]], np.float32)
    pos[0, axis] = np.float32(bound - side * 1e-4)
    vel = np.zeros((1, 3), np.float32)
    vel[0, axis] = 10.0 * side
    ev = np.zeros((1, 4), np.float32)
    po_, vo, eo = po.integrate_f64(p, pos, vel, ev, np.zeros((1, 3)), np.array([1000.0]))
    assert po_[0, axis] == bound - side * kern
    assert vo[0, axis] * side < 0                                 # reflected
    others = [a for a in range(3) if a != axis]
    assert 4 * vo[0, axis] ** 2 + sum(vo[0, a] ** 2 for a in others) == pytest.approx(36.0, rel=1e-13)
    assert abs(vo[0, axis]) == pytest.approx(3.0, rel=1e-6)


def test_pair_stats_known_answers(po):
    """getSmsTaskPairStats (sph_arrangement.cu:1595-1636) on hand-counted
    task lists.  The 70-particle cell of test_arrange_tasks_known_answer has
    3 tasks + pad: pair (0,1) is one cell, second task holds 70 - 32 = 38 >= 32
    particles (full), first.xxi 0 != second.xxx 2 (widened); the odd-tail
    pair (2, pad) is one cell, the pad holds 70 - 96 < 32 (partial) and
    copies task 2's x bounds 2..2 (compact)."""
    p = make_params(h=0.04, world=(0.12, 0.12, 0.12))
    rows = []
    for (mx, my, mz), k in [((0, 0, 0), 20), ((0, 2, 1), 20), ((2, 3, 3), 30)]:
        for j in range(k):
            rows.append([0.04 + 0.01 * mx + 0.0005 + 1e-5 * j, 0.04 + 0.01 * my + 0.005,
                         0.04 + 0.01 * mz + 0.005, 0])
    pos = np.array(rows, np.float32)[::-1].copy()
    g = po.arrange(p, pos)
    st = po.pair_stats(g["tasks"], g["tasks_count"], g["count"])
    assert st == dict(total_pairs=2, same_cell_pairs=2, is_same_pairs=2, full_is_same_pairs=1,
                      partial_is_same_pairs=1, compact_is_same_pairs=1, widened_is_same_pairs=1)
    # 64 particles in micro x layer 1: two tasks, 64 - 32 = 32 (full), x bounds
    # 1..1 both (compact)
    rows = [[0.04 + 0.015, 0.04 + 0.005 + 0.01 * (j % 4), 0.04 + 0.01 * (j // 16) + 0.005, 0]
            for j in range(64)]
    g = po.arrange(p, np.array(rows, np.float32))
    st = po.pair_stats(g["tasks"], g["tasks_count"], g["count"])
    assert st == dict(total_pairs=1, same_cell_pairs=1, is_same_pairs=1, full_is_same_pairs=1,
                      partial_is_same_pairs=0, compact_is_same_pairs=1, widened_is_same_pairs=0)
    # two cells of 30 particles each ((30 + 27) >> 5 = 1 task each, no
    # neighbours): one pair of two cells, not isSame
    p3 = make_params(h=0.04, world=(0.24, 0.12, 0.12))
    rng = np.random.default_rng(3)
    a = np.zeros((60, 4), np.float32)
    a[:30, :3] = rng.uniform([0.042, 0.042, 0.042], [0.078, 0.078, 0.078], (30, 3))
    a[30:, :3] = rng.uniform([0.162, 0.042, 0.042], [0.198, 0.078, 0.078], (30, 3))
    g = po.arrange(p3, a)
    assert g["tasks_count"] == 2
    st = po.pair_stats(g["tasks"], g["tasks_count"], g["count"])
    assert st == dict(total_pairs=1, same_cell_pairs=0, is_same_pairs=0, full_is_same_pairs=0,
                      partial_is_same_pairs=0, compact_is_same_pairs=0, widened_is_same_pairs=0)
    # no tasks: all zero
    st = po.pair_stats(g["tasks"][:0], 0, g["count"])
    assert all(v == 0 for v in st.values())


def test_checksum_definition(po):
    pos = np.arange(2048 * 4, dtype=np.float32).reshape(2048, 4) * np.float32(1e-3)
    out = po.checksum(pos)
    d = pos.astype(np.float64)
    w = (np.arange(2048) & 1023) + 1
    exp = [d[:, 0].sum(), d[:, 1].sum(), d[:, 2].sum(), d[:, 3].sum(),
           (w * (d[:, 0] + 3 * d[:, 1] + 7 * d[:, 2])).sum()]
    np.testing.assert_allclose(out, exp, rtol=1e-14)


def test_f32_step_tracks_fp64(po):
    """The FP32 CPU baseline agrees with the FP64 oracle to ~1e-6."""
    p, pos = cfg0_params(po)
    n = len(pos)
    r = po.fp64_step(p, pos, np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32))
    P, V, E, A = po.step_f32(p, pos, nsteps=1)
    assert (np.abs(P[:, 3] - 1 / r["rho"]) * r["rho"]).max() < 2e-6
    err = np.linalg.norm(A - r["acc"], axis=1) / np.maximum(np.linalg.norm(r["acc"], axis=1), r["acc_scale"])
    assert err.max() < 5e-6
