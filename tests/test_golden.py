"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from the
oracle): the oracle must keep reproducing them exactly, and the GPU path must
match their arrange arrays bit for bit and their FP64 step to 1e-5."""
from pathlib import Path

import numpy as np
import pytest

from tolerance import REL_TOL, step_errors

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def load(name):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))


def oparams(po, arr):
    p = po.OrcParams()
    v = [float(x) for x in arr]
    p.kernel_h, p.mass, p.rest_density, p.gas_constant, p.viscosity, p.time_step = v[:6]
    p.gravity[0], p.gravity[1], p.gravity[2] = v[6:9]
    p.wall_damping, p.surface_normal, p.surface_coe = v[9:12]
    p.world_size[0], p.world_size[1], p.world_size[2] = v[12:15]
    return p


def test_fixture_set_complete():
    assert len(NAMES) >= 11


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(po, name):
    f = load(name)
    p = oparams(po, f["params"])
    g = po.arrange(p, f["pos"])
    for k in ("hash", "perm", "begin", "count", "req", "compact"):
        assert np.array_equal(g[k], f[k]), k
    assert g["middle_raw"] == f["middle_raw"] and g["tasks_count"] == f["tasks_count"]
    if len(g["tasks"]):
        assert np.array_equal(g["tasks"].view(np.uint8).reshape(-1, 32), f["tasks"])
    n = len(f["pos"])
    r = po.fp64_step(p, f["pos"], np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32))
    for k in ("rho", "pressure", "acc", "acc_scale"):
        np.testing.assert_array_equal(r[k], f[k])
    np.testing.assert_array_equal(r["pos"], f["pos1"])


def test_fixture_split_known_answers():
    # cells of exactly 14/15/33/36/37 particles (neighbour totals < 60)
    expect = {"dense_cell_14": (0, -1), "dense_cell_15": (1, 0), "dense_cell_33": (1, 1),
              "dense_cell_36": (1, 4), "dense_cell_37": (2, 0)}
    for name, (req, tra) in expect.items():
        f = load(name)
        assert f["req"].max() == req, name
        assert f["tasks_count"] == req
        assert f["middle_raw"] == tra, name


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_fixture(gsph, po, name):
    from helpers import to_gsph_params
    f = load(name)
    op = oparams(po, f["params"])
    print(f"\n[gpu-case] golden {name}", flush=True)
    s = gsph.System(to_gsph_params(gsph, op), f["pos"])
    s.arrange()
    g = s.read_grid()
    for k in ("hash", "perm", "begin", "count", "req", "compact"):
        assert np.array_equal(g[k], f[k]), k
    assert g["middle_raw"] == f["middle_raw"] and g["tasks"].shape[0] == f["tasks"].shape[0]
    # every task descriptor byte for byte, the odd-tail pad included
    gt = g["tasks"].view(np.uint8).reshape(-1, 32)
    assert np.array_equal(gt, f["tasks"])
    # the seven pair statistics against the oracle's over the fixture's tasks
    T = int(f["tasks_count"])
    exp = po.pair_stats(f["tasks"].reshape(-1).view(po.TASK_DTYPE), T, f["count"])
    assert s.pair_stats() == exp
    s.density()
    s.force()
    s.integrate()
    pos1, vel1, ev1, acc1 = s.read_state()
    ref = dict(rho=f["rho"], pressure=f["pressure"], acc=f["acc"], acc_scale=f["acc_scale"],
               pos=f["pos1"], vel=f["vel1"])
    for k, v in step_errors(pos1, vel1, ev1, acc1, ref, op).items():
        assert v <= REL_TOL, f"{k} {v:.3e}"
    s.close()
