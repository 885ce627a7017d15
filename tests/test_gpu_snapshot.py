"""Checkpoint / resume on the GPU: a context restored from a snapshot
continues bit-exactly, for a plain scene and for a batch of scenes."""
import numpy as np
import pytest

from helpers import mixed_scene, to_gsph_params

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"\n[gpu-case] {msg}", flush=True)


def test_snapshot_resume_plain(gsph, po, tmp_path):
    op, pos = mixed_scene(po)
    banner(f"snapshot plain n={len(pos)}")
    s = gsph.System(to_gsph_params(gsph, op), pos)
    s.step(4)
    path = tmp_path / "plain.gsnap"
    s.save(path)
    assert gsph.snapshot_check(path) == (len(pos), 1, 4)
    s.step(5)
    ref = s.read_state()
    s.close()
    r = gsph.load_snapshot(path)
    assert r.step_count() == 4 and r.n == len(pos)
    r.step(5)
    got = r.read_state()
    assert r.step_count() == 9
    r.close()
    for k in range(4):
        assert np.array_equal(ref[k], got[k])


def test_snapshot_resume_batch(gsph, po, tmp_path):
    worlds, blocks = [], []
    for k, (nx, ny, nz) in enumerate([(10, 12, 9), (16, 6, 8), (7, 7, 7)]):
        blocks.append(po.scene_block([0.06, 0.06, 0.06], nx, ny, nz, 0.01, 0.01, 40 + k))
        worlds.append([2 * nx * 0.01 + 0.12, 2 * ny * 0.01 + 0.12, 2 * nz * 0.01 + 0.12])
    counts = [len(b) for b in blocks]
    banner(f"snapshot batch counts={counts}")
    p = gsph.params_default()
    p.time_step = 1e-4
    s = gsph.System(p, np.concatenate(blocks), batch_worlds=worlds, batch_counts=counts)
    s.step(3)
    path = tmp_path / "batch.gsnap"
    s.save(path)
    s.step(4)
    ref = s.read_state()
    s.close()
    assert gsph.snapshot_check(path) == (sum(counts), 3, 3)
    r = gsph.load_snapshot(path)
    r.step(4)
    got = r.read_state()
    r.close()
    for k in range(4):
        assert np.array_equal(ref[k], got[k])
