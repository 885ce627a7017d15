"""Serving form: continuous batching gives every request exactly the result
it gets alone, under admission by budget and with failing requests isolated."""
import numpy as np
import pytest

from helpers import make_params, to_gsph_params

pytestmark = pytest.mark.gpu


def request_scenes(po, k=6, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        dims = [int(x) for x in rng.integers(5, 16, 3)]
        spacing = 0.005 if i == 3 else 0.01        # one dense (SMS) request
        pos = po.scene_block([0.06, 0.06, 0.06], *dims, spacing, 0.01, 300 + i)
        ext = np.array(dims) * spacing
        world = [float(max(0.24, 2 * e + 0.12)) for e in ext]
        steps = int(rng.integers(3, 9))
        arrival = float(rng.integers(0, 6))
        out.append((world, pos, steps, arrival))
    return out


def alone(gsph, op, world, pos, steps, vel=None):
    gp = to_gsph_params(gsph, op)
    for a in range(3):
        gp.world_size[a] = world[a]
    s = gsph.System(gp, pos, vel)
    s.step(steps)
    p, v, e, _ = s.read_state()
    s.close()
    return p, v, e


def test_results_same_alone_and_in_the_mix(gsph, po):
    op = make_params(h=0.02, world=(1, 1, 1))
    reqs = request_scenes(po)
    print("\n[gpu-case] serving mix", [len(r[1]) for r in reqs], flush=True)
    srv = gsph.Server(to_gsph_params(gsph, op), hbm_budget_bytes=8e9, max_scenes=4)
    ids = [srv.submit(w, p, st, arr) for (w, p, st, arr) in reqs]
    srv.run(realtime=False)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["max_active"] > 1 and st["recompositions"] > 1    # requests overlapped
    for rid, (w, p, steps, arr) in zip(ids, reqs):
        r = srv.result(rid)
        assert r["state"] == "done" and r["latency"] >= steps
        ap, av, ae = alone(gsph, op, w, p, steps)
        assert np.array_equal(r["pos"], ap)
        assert np.array_equal(r["vel"], av)
        assert np.array_equal(r["ev"], ae)
    srv.close()


def test_budget_defers_and_rejects(gsph, po):
    op = make_params(h=0.02, world=(1, 1, 1))
    reqs = request_scenes(po, k=4, seed=8)
    # the library's admission estimate: 296 B + the default SMS mask budget (1024 B)
    # per particle, 24 B per cell, 1 MiB
    def est(w, p):
        cells = np.prod([int(np.ceil(np.float32(x) / np.float32(0.02))) for x in w])
        return len(p) * (296 + 1024) + cells * 24 + (1 << 20)
    sizes = [est(w, p) for (w, p, _, _) in reqs]
    s2 = sorted(sizes)
    need = s2[0] + s2[1] - 1                     # never two requests at once ...
    assert need >= s2[-1]                        # ... but each one alone fits
    srv = gsph.Server(to_gsph_params(gsph, op), hbm_budget_bytes=need, max_scenes=64)
    ids = [srv.submit(w, p, st, 0.0) for (w, p, st, _) in reqs]
    huge = po.scene_block([0.06, 0.06, 0.06], 30, 30, 30, 0.01, 0.0, 1)
    big_id = srv.submit([0.72, 0.72, 0.72], huge, 2, 0.0)
    srv.run(realtime=False)
    st = srv.stats()
    assert st["rejected"] == 1 and srv.result(big_id)["state"] == "rejected"
    assert st["completed"] == len(reqs)
    assert st["peak_bytes"] <= need and st["max_active"] == 1
    for rid, (w, p, steps, _) in zip(ids, reqs):
        ap, _, _ = alone(gsph, op, w, p, steps)
        assert np.array_equal(srv.result(rid)["pos"], ap)
    srv.close()


def test_failing_request_is_isolated(gsph, po):
    op = make_params(h=0.02, world=(1, 1, 1))
    reqs = request_scenes(po, k=3, seed=11)
    srv = gsph.Server(to_gsph_params(gsph, op), hbm_budget_bytes=8e9, max_scenes=8)
    ids = []
    for i, (w, p, st, _) in enumerate(reqs):
        vel = None
        if i == 1:                      # a NaN velocity poisons this scene only
            vel = np.zeros((len(p), 3), np.float32)
            vel[0, 0] = np.nan
        ids.append(srv.submit(w, p, 5, 0.0, vel))
    bad_box = reqs[0][1].copy()
    bad_box[0, 0] = 5.0                 # outside its box: rejected at submission
    rej = srv.submit(reqs[0][0], bad_box, 3, 0.0)
    srv.run(realtime=False)
    assert srv.result(rej)["state"] == "rejected"
    assert srv.result(ids[1])["state"] == "failed"
    for i in (0, 2):
        w, p, _, _ = reqs[i]
        r = srv.result(ids[i])
        assert r["state"] == "done"
        ap, _, _ = alone(gsph, op, w, p, 5)
        assert np.array_equal(r["pos"], ap)
    srv.close()
