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
    """A budget that holds an engine for any one request but not for two:
    requests wait for each other (never two resident), the peak of the
    engine's allocated bytes (both engines while one replaces the other)
    stays within the budget, an oversized request is rejected, and every
    result equals the request's solo run."""
    op = make_params(h=0.02, world=(1, 1, 1))
    rng = np.random.default_rng(8)
    reqs = []
    for i in range(4):
        dims = [int(x) for x in rng.integers(28, 34, 3)]
        pos = po.scene_block([0.06, 0.06, 0.06], *dims, 0.01, 0.01, 400 + i)
        world = [float(2 * d * 0.01 + 0.12) for d in dims]
        reqs.append((world, pos, int(rng.integers(3, 7))))
    gp = to_gsph_params(gsph, op)
    opt = gsph.options_default()
    cells = [int(np.prod([int(np.ceil(np.float32(x) / np.float32(0.02))) for x in w]))
             for (w, _, _) in reqs]
    alone_bytes = [gsph.engine_bytes(len(p), c, 1, opt) for (_, p, _), c in zip(reqs, cells)]
    budget = 1.3 * max(alone_bytes)
    # two requests together never fit, nor does an engine grown for two
    assert min(alone_bytes) + min(alone_bytes) > budget
    srv = gsph.Server(gp, hbm_budget_bytes=budget, max_scenes=64, options=opt)
    ids = [srv.submit(w, p, st, 0.0) for (w, p, st) in reqs]
    huge = po.scene_block([0.06, 0.06, 0.06], 60, 60, 60, 0.01, 0.0, 1)
    big_id = srv.submit([1.32, 1.32, 1.32], huge, 2, 0.0)
    print(f"\n[gpu-case] serving budget {budget:.3e}", flush=True)
    srv.run(realtime=False)
    st = srv.stats()
    assert st["rejected"] == 1 and srv.result(big_id)["state"] == "rejected"
    assert st["completed"] == len(reqs)
    assert st["peak_bytes"] <= budget and st["max_active"] == 1
    assert max(srv.result(i)["queued"] for i in ids) > 0        # the budget made them wait
    for rid, (w, p, steps) in zip(ids, reqs):
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


def test_config4_serving_workload(gsph, po):
    """The serving workload (Poisson arrivals; N log-uniform in 8,192 ..
    1,048,576; 50..500 steps) on 12 requests, with an HBM budget that binds
    (2.5x the engine of the largest request alone): everything completes,
    requests waited for memory, the engine's peak allocation stays within the
    budget, and sampled requests (the smallest, the largest, two more) give
    bitwise their solo runs."""
    params, _ = gsph.scene_config(0)
    worlds, blocks, steps, rng = gsph.serving_workload(12, seed=20261207)
    arrivals = np.cumsum(rng.exponential(40.0, len(blocks)))   # engine iterations
    opt = gsph.options_default()
    ns = [len(b) for b in blocks]
    cells = [int(np.prod([int(np.ceil(np.float32(x) / np.float32(0.02))) for x in w]))
             for w in worlds]
    alone_bytes = [gsph.engine_bytes(n, c, 1, opt) for n, c in zip(ns, cells)]
    budget = 2.5 * max(alone_bytes)
    assert sum(alone_bytes) > budget
    print(f"\n[gpu-case] config4 serving n={sum(ns)} budget={budget:.3e}", flush=True)
    srv = gsph.Server(params, hbm_budget_bytes=budget, max_scenes=4096, options=opt)
    ids = [srv.submit(w, b, st, float(a)) for w, b, st, a in zip(worlds, blocks, steps, arrivals)]
    srv.run(realtime=False)
    st = srv.stats()
    print(st, flush=True)
    assert st["completed"] == len(ids) and st["failed"] == 0 and st["rejected"] == 0
    assert st["peak_bytes"] <= budget and st["max_active"] > 1
    assert max(srv.result(i)["queued"] for i in ids) > 0
    order = np.argsort(ns)
    sample = sorted({int(order[0]), int(order[-1]), int(order[len(order) // 2]), 3})
    for k in sample:
        p = gsph.Params.from_buffer_copy(params)
        for a in range(3):
            p.world_size[a] = worlds[k][a]
        s = gsph.System(p, blocks[k])
        s.step(steps[k])
        ap, av, ae, _ = s.read_state()
        s.close()
        r = srv.result(ids[k])
        assert r["state"] == "done"
        assert np.array_equal(r["pos"], ap), k
        assert np.array_equal(r["vel"], av), k
        assert np.array_equal(r["ev"], ae), k
    srv.close()
