"""Shared test inputs: parameter blocks and synthetic scenes (seeded)."""
import numpy as np


def make_params(po=None, h=0.02, mass=1e-3, world=(0.64, 0.64, 0.64), dt=1e-4, g=-9.81):
    import pyoracle

    p = pyoracle.OrcParams()
    p.kernel_h = h
    p.mass = mass
    p.rest_density = 1000.0
    p.gas_constant = 1.0
    p.viscosity = 6.5
    p.time_step = dt
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, g, 0.0
    p.wall_damping = -0.5
    p.surface_normal = 0.1
    p.surface_coe = 0.2
    for a in range(3):
        p.world_size[a] = world[a]
    return p


def cfg0_params(po):
    """Workload 0: 32^3 jittered block, spacing 0.01, box 0.64^3, seed 20261203."""
    p = make_params(h=0.02, mass=1e-3, world=(0.64, 0.64, 0.64))
    pos = po.scene_block([0.06, 0.06, 0.06], 32, 32, 32, 0.01, 0.01, 20261203)
    return p, pos


def to_gsph_params(gsph, op):
    """gsph.Params with the same values as an oracle parameter block."""
    p = gsph.Params()
    for name, _ in gsph.Params._fields_:
        v = getattr(op, name)
        if hasattr(v, "__len__"):
            for i in range(3):
                getattr(p, name)[i] = v[i]
        else:
            setattr(p, name, v)
    return p


def dense_scene(po, seed=1, n_side=16, spacing=0.005, h=0.02, world=0.24):
    """Dense block (about 64 particles per cell): exercises the SMS path."""
    p = make_params(h=h, world=(world, world, world))
    pos = po.scene_block([0.05, 0.05, 0.05], n_side, n_side, n_side, spacing, 0.01, seed)
    return p, pos


def mixed_scene(po, seed=2):
    """A lattice block at spacing h/2 (TRA) next to a fine blob (SMS), with
    cells of every occupancy in between, in a non-cubic box."""
    p = make_params(h=0.02, world=(0.36, 0.24, 0.30))
    a = po.scene_block([0.05, 0.05, 0.05], 12, 8, 10, 0.01, 0.01, seed)
    rng = np.random.default_rng(seed)
    blob = np.zeros((6000, 4), np.float32)
    blob[:, :3] = np.clip(rng.normal([0.25, 0.12, 0.18], [0.025, 0.02, 0.03], (6000, 3)),
                          [0.031, 0.031, 0.031], [0.329, 0.209, 0.269])
    sparse = np.zeros((800, 4), np.float32)
    sparse[:, :3] = rng.uniform([0.031] * 3, [0.329, 0.209, 0.269], (800, 3))
    return p, np.concatenate([a, blob, sparse])


def skewed_scene(po, seed=5):
    """Small analogue of the skewed-occupancy workload: a pool at spacing
    h/2 plus a sphere at spacing h/4 (8x denser) above it."""
    p = make_params(h=0.02, world=(0.36, 0.30, 0.36))
    pool = po.scene_block([0.04, 0.04, 0.04], 28, 6, 28, 0.01, 0.01, seed)
    s, r, c = 0.005, 0.045, np.array([0.18, 0.19, 0.18])
    m = int(np.ceil(2 * r / s))
    g = (np.arange(m) + 0.5) * s - r
    xx, yy, zz = np.meshgrid(g, g, g, indexing="ij")
    keep = xx ** 2 + yy ** 2 + zz ** 2 < r * r
    rng = np.random.default_rng(seed)
    pts = np.stack([xx[keep], yy[keep], zz[keep]], 1) + c
    pts += rng.uniform(-0.01 * s, 0.01 * s, pts.shape)
    sphere = np.zeros((len(pts), 4), np.float32)
    sphere[:, :3] = pts
    return p, np.concatenate([pool, sphere])

