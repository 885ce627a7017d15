"""Particle insertion between steps (gsph_add_particles) on the GPU.

A context that grows by appending particles must continue exactly like a
context created from the same state: the next arrange sorts every particle,
and the buffers that grow (capacity doubling) keep the state.  One step after
an insertion is also checked against the FP64 oracle."""
import numpy as np
import pytest

from helpers import mixed_scene, to_gsph_params
from tolerance import REL_TOL, step_errors

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"\n[gpu-case] {msg}", flush=True)


def fresh_from(gsph, params, state, extra_pos=None, extra_vel=None, reset_all=False):
    """A new context holding `state` plus appended particles, built with
    gsph_create + gsph_set_state (the path the insertion must reproduce)."""
    pos, vel, ev = state[0].copy(), state[1].copy(), state[2].copy()
    if reset_all:
        ev[:] = 0
    if extra_pos is not None:
        p = extra_pos.copy()
        p[:, 3] = 0
        pos = np.concatenate([pos, p])
        vel = np.concatenate([vel, extra_vel if extra_vel is not None
                              else np.zeros((len(p), 3), np.float32)])
        ev = np.concatenate([ev, np.zeros((len(p), 4), np.float32)])
    s = gsph.System(params, pos, vel)
    s.set_state(pos, vel, ev)
    return s


def test_insertion_grows_and_matches_fresh_context(gsph, po):
    op, pos = mixed_scene(po)
    params = to_gsph_params(gsph, op)
    n = len(pos)
    rng = np.random.default_rng(11)
    vel_new = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    cuts = [4000, 6000, n]
    banner(f"insert mixed n={n} cuts={cuts}")
    s = gsph.System(params, pos[:cuts[0]])
    r = gsph.System(params, pos[:cuts[0]])
    s.step(2)
    r.step(2)
    for a, b in zip(cuts[:-1], cuts[1:]):
        # the first insertion exceeds the capacity (grows to 8000), the
        # second fits in it
        s.add_particles(pos[a:b], vel_new[a:b])
        st = r.read_state()
        r.close()
        r = fresh_from(gsph, params, st, pos[a:b], vel_new[a:b])
        assert s.n == r.n == b
        s.step(2)
        r.step(2)
        got, ref = s.read_state(), r.read_state()
        for k in range(4):
            assert np.array_equal(got[k], ref[k]), (b, k)
        assert np.array_equal(s.checksum(), r.checksum())
    s.close()
    r.close()


def test_insertion_step_matches_oracle(gsph, po):
    op, pos = mixed_scene(po)
    params = to_gsph_params(gsph, op)
    n = len(pos)
    banner(f"insert oracle n={n}")
    s = gsph.System(params, pos[:5000])
    s.step(1)
    vel_new = np.tile(np.array([[0.0, -0.3, 0.1]], np.float32), (n - 5000, 1))
    s.add_particles(pos[5000:], vel_new)
    pos0, vel0, ev0, _ = s.read_state()
    assert np.all(pos0[5000:, 3] == 0) and np.all(ev0[5000:] == 0)
    assert np.array_equal(vel0[5000:], vel_new)
    s.step(1)
    pos1, vel1, ev1, acc1 = s.read_state()
    s.close()
    ref = po.fp64_step(op, pos0, vel0, ev0)
    errs = step_errors(pos1, vel1, ev1, acc1, ref, op)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= REL_TOL, f"{k} error {v:.3e}"


def test_insertion_reset_all_and_graph(gsph, po):
    """reset_all zeroes ev of the particles already there; a captured step
    graph is dropped and captured again for the new size."""
    op, pos = mixed_scene(po)
    params = to_gsph_params(gsph, op)
    banner("insert reset_all + graph")
    s = gsph.System(params, pos[:6000])
    s.enable_graph(True)
    s.step(2)
    st = s.read_state()
    s.add_particles(pos[6000:], reset_all=True)
    s.step(2)
    r = fresh_from(gsph, params, st, pos[6000:], reset_all=True)
    r.step(2)
    got, ref = s.read_state(), r.read_state()
    for k in range(4):
        assert np.array_equal(got[k], ref[k]), k
    s.close()
    r.close()


def test_insertion_rejects_bad_input(gsph, po):
    op, pos = mixed_scene(po)
    params = to_gsph_params(gsph, op)
    banner("insert errors")
    s = gsph.System(params, pos[:3000])
    s.step(1)
    before = s.read_state()
    bad = pos[3000:3010].copy()
    bad[4, 1] = -0.01
    with pytest.raises(gsph.GsphError) as e:
        s.add_particles(bad)
    assert e.value.code == -3          # GSPH_ERR_OUT_OF_DOMAIN
    bad[4, 1] = np.nan
    with pytest.raises(gsph.GsphError):
        s.add_particles(bad)
    s.add_particles(pos[:0])           # nothing to add
    assert s.n == 3000
    after = s.read_state()
    for k in range(3):
        assert np.array_equal(before[k], after[k])
    s.step(1)                          # the context is still usable
    b = gsph.System(params, np.concatenate([pos[:100], pos[200:300]]),
                    batch_worlds=[list(op.world_size), list(op.world_size)], batch_counts=[100, 100])
    with pytest.raises(gsph.GsphError) as e:
        b.add_particles(pos[300:310])
    assert e.value.code == -1          # GSPH_ERR_INVALID_ARG
    b.close()
    s.close()
