"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_build" / "liboracle.so"


class OrcParams(C.Structure):
    _fields_ = [
        ("kernel_h", C.c_float), ("mass", C.c_float), ("rest_density", C.c_float),
        ("gas_constant", C.c_float), ("viscosity", C.c_float), ("time_step", C.c_float),
        ("gravity", C.c_float * 3), ("wall_damping", C.c_float),
        ("surface_normal", C.c_float), ("surface_coe", C.c_float),
        ("world_size", C.c_float * 3),
    ]


class OrcConstants(C.Structure):
    _fields_ = [
        ("kernel_h", C.c_float), ("kernel_2", C.c_float), ("cell_size", C.c_float),
        ("inv_cell_size", C.c_float), ("grid_size", C.c_int * 3),
        ("mass", C.c_float), ("rest_density", C.c_float), ("gas_constant", C.c_float),
        ("viscosity", C.c_float), ("time_step", C.c_float), ("gravity", C.c_float * 3),
        ("wall_damping", C.c_float), ("surface_normal", C.c_float), ("surface_coe", C.c_float),
        ("poly6_value", C.c_float), ("spiky_value", C.c_float), ("visco_value", C.c_float),
        ("grad_poly6", C.c_float), ("lplc_poly6", C.c_float), ("self_density", C.c_float),
        ("inv_rest_density", C.c_float), ("density_scale", C.c_float),
        ("half_spiky_value", C.c_float), ("viscosity_visco_value", C.c_float),
        ("grad_color_scale", C.c_float), ("lplc_color_scale", C.c_float),
        ("bound_min", C.c_float * 3), ("bound_max", C.c_float * 3),
        ("world_size", C.c_float * 3),
    ]


TASK_DTYPE = np.dtype([
    ("is_same", "i1"), ("pad0", "u1"), ("paired_bounds", "<u2"), ("cellid", "<i4"),
    ("cell_pos", "<u2", (3,)), ("p_offset", "<u2"),
    ("xxi", "i1"), ("xxx", "i1"), ("yyi", "i1"), ("yyy", "i1"), ("zzi", "i1"), ("zzz", "i1"),
    ("pad1", "i1", (2,)), ("cell_begin", "<i4"), ("cell_particle_count", "<i4"),
])

_P = C.c_void_p
_I = C.c_int
_SIGS = {
    "orc_derive": (_I, [C.POINTER(OrcParams), C.POINTER(OrcConstants)]),
    "orc_splitmix64": (C.c_uint64, [C.POINTER(C.c_uint64)]),
    "orc_scene_block": (_I, [_P, _I, _I, _I, C.c_float, C.c_float, C.c_uint64, _P]),
    "orc_scene_lattice": (_I, [_P, _P, C.c_float, C.c_float, _P, _I]),
    "orc_hash": (_I, [C.POINTER(OrcParams), _I, _P, _P]),
    "orc_split_req": (_I, [_P, _P, _P]),
    "orc_arrange": (_I, [C.POINTER(OrcParams), _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I]),
    "orc_density_f64": (_I, [C.POINTER(OrcParams), _I, _P, _P, _P]),
    "orc_density_brute_f64": (_I, [C.POINTER(OrcParams), _I, _P, _P]),
    "orc_force_f64": (_I, [C.POINTER(OrcParams), _I, _P, _P, _P, _P, _P, _P, _P]),
    "orc_integrate_f64": (_I, [C.POINTER(OrcParams), _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "orc_step_f32": (_I, [C.POINTER(OrcParams), _I, _P, _P, _P, _P, _I, _I]),
    "orc_checksum": (_I, [_I, _P, _P]),
    "orc_pair_stats_of": (_I, [_P, _I, _P, _P]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} missing; run `make -C oracle`")
        lib = C.CDLL(str(LIB_PATH))
        for k, (r, a) in _SIGS.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def params_from(obj) -> OrcParams:
    """OrcParams from any object with the gsph_params fields (or a dict)."""
    get = (lambda k: obj[k]) if isinstance(obj, dict) else (lambda k: getattr(obj, k))
    p = OrcParams()
    for name, _ in OrcParams._fields_:
        v = get(name)
        if hasattr(v, "__len__"):
            for i in range(3):
                getattr(p, name)[i] = float(v[i])
        else:
            setattr(p, name, float(v))
    return p


def derive(p: OrcParams) -> OrcConstants:
    c = OrcConstants()
    load().orc_derive(C.byref(p), C.byref(c))
    return c


def splitmix64(state: int, count: int):
    st = C.c_uint64(state)
    return [load().orc_splitmix64(C.byref(st)) for _ in range(count)]


def scene_block(corner, nx, ny, nz, spacing, jitter_frac, seed):
    out = np.zeros((nx * ny * nz, 4), np.float32)
    load().orc_scene_block(_p(np.asarray(corner, np.float32)), nx, ny, nz, spacing,
                           jitter_frac, seed, _p(out))
    return out


def scene_lattice(begin, end, h, interval):
    b = np.asarray(begin, np.float32)
    e = np.asarray(end, np.float32)
    n = load().orc_scene_lattice(_p(b), _p(e), h, interval, None, 0)
    out = np.zeros((n, 4), np.float32)
    load().orc_scene_lattice(_p(b), _p(e), h, interval, _p(out), n)
    return out


def hash_(p: OrcParams, pos4):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    key = np.zeros(pos4.shape[0], np.int32)
    bad = load().orc_hash(C.byref(p), pos4.shape[0], _p(pos4), _p(key))
    return key, bad


def split_req(count, grid):
    count = np.ascontiguousarray(count, np.int32)
    g = np.asarray(grid, np.int32)
    req = np.zeros_like(count)
    load().orc_split_req(_p(count), _p(g), _p(req))
    return req


def arrange(p: OrcParams, pos4):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    n = pos4.shape[0]
    c = derive(p)
    ncell = c.grid_size[0] * c.grid_size[1] * c.grid_size[2]
    out = dict(hash=np.zeros(n, np.int32), perm=np.zeros(n, np.int32),
               begin=np.zeros(ncell, np.int32), count=np.zeros(ncell, np.int32),
               req=np.zeros(ncell, np.int32), compact=np.zeros(n, np.int32))
    mid = C.c_int()
    T = C.c_int()
    cap = (n + 31) // 32 + min(ncell, n) + 2
    tasks = np.zeros(cap, TASK_DTYPE)
    rc = load().orc_arrange(C.byref(p), n, _p(pos4), _p(out["hash"]), _p(out["perm"]),
                            _p(out["begin"]), _p(out["count"]), _p(out["req"]),
                            _p(out["compact"]), C.byref(mid), C.byref(T), _p(tasks), cap)
    if rc != 0:
        raise RuntimeError(f"orc_arrange failed: {rc}")
    out["middle_raw"] = mid.value
    out["tasks_count"] = T.value
    out["tasks"] = tasks[:T.value + (T.value & 1)]
    return out


PAIR_STATS_FIELDS = ("total_pairs", "same_cell_pairs", "is_same_pairs", "full_is_same_pairs",
                     "partial_is_same_pairs", "compact_is_same_pairs", "widened_is_same_pairs")


def pair_stats(tasks, num_tasks, count):
    """Arrangement::getSmsTaskPairStats over a task array (T tasks plus the
    odd-tail pad) and the per-cell counts; a dict of the seven fields."""
    tasks = np.ascontiguousarray(tasks, TASK_DTYPE)
    if num_tasks > 0 and len(tasks) < num_tasks + (num_tasks & 1):
        raise ValueError("task array shorter than T plus the pad")
    count = np.ascontiguousarray(count, np.int32)
    out = (C.c_int * 7)()
    load().orc_pair_stats_of(_p(tasks), int(num_tasks), _p(count), C.cast(out, C.c_void_p))
    return dict(zip(PAIR_STATS_FIELDS, list(out)))


def density_f64(p, pos4):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    n = pos4.shape[0]
    rho = np.zeros(n)
    pr = np.zeros(n)
    load().orc_density_f64(C.byref(p), n, _p(pos4), _p(rho), _p(pr))
    return rho, pr


def density_brute_f64(p, pos4):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    rho = np.zeros(pos4.shape[0])
    load().orc_density_brute_f64(C.byref(p), pos4.shape[0], _p(pos4), _p(rho))
    return rho


def force_f64(p, pos4, ev4, rho, pressure, alt=False):
    """FP64 acceleration and its scale; with alt=True also the acceleration of
    the other surface-force branch (equal to acc except where |grad c| is
    within FP32 rounding of the threshold)."""
    pos4 = np.ascontiguousarray(pos4, np.float32)
    ev4 = np.ascontiguousarray(ev4, np.float32)
    n = pos4.shape[0]
    acc = np.zeros((n, 3))
    scale = np.zeros(n)
    acc_alt = np.zeros((n, 3)) if alt else None
    load().orc_force_f64(C.byref(p), n, _p(pos4), _p(ev4), _p(np.ascontiguousarray(rho)),
                         _p(np.ascontiguousarray(pressure)), _p(acc), _p(scale), _p(acc_alt))
    return (acc, scale, acc_alt) if alt else (acc, scale)


def integrate_f64(p, pos4, vel3, ev4, acc, rho):
    n = pos4.shape[0]
    po = np.zeros((n, 4))
    vo = np.zeros((n, 3))
    eo = np.zeros((n, 4))
    load().orc_integrate_f64(C.byref(p), n, _p(np.ascontiguousarray(pos4, np.float32)),
                             _p(np.ascontiguousarray(vel3, np.float32)),
                             _p(np.ascontiguousarray(ev4, np.float32)),
                             _p(np.ascontiguousarray(acc, np.float64)),
                             _p(np.ascontiguousarray(rho, np.float64)), _p(po), _p(vo), _p(eo))
    return po, vo, eo


def step_f32(p, pos4, vel3=None, ev4=None, nsteps=1, nthreads=0):
    """FP32 reference step in place on copies; returns (pos, vel, ev, acc)."""
    pos = np.array(pos4, np.float32, copy=True).reshape(-1, 4)
    n = pos.shape[0]
    vel = np.zeros((n, 3), np.float32) if vel3 is None else np.array(vel3, np.float32)
    ev = np.zeros((n, 4), np.float32) if ev4 is None else np.array(ev4, np.float32)
    acc = np.zeros((n, 3), np.float32)
    rc = load().orc_step_f32(C.byref(p), n, _p(pos), _p(vel), _p(ev), _p(acc), nsteps, nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_step_f32 failed: {rc}")
    return pos, vel, ev, acc


def checksum(pos4):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    out = np.zeros(5)
    load().orc_checksum(pos4.shape[0], _p(pos4), _p(out))
    return out


def fp64_step(p, pos4, vel3, ev4):
    """One FP64 step from an FP32 state (the reference's tick in FP64).

    Arrange with the oracle (bit-exact) to get the sorted order, then density,
    force and integration in FP64 on the sorted state.  Returns a dict with the
    sorted FP32 inputs and the FP64 results.
    """
    g = arrange(p, pos4)
    perm = g["perm"]
    pos_s = np.ascontiguousarray(pos4[perm])
    vel_s = np.ascontiguousarray(vel3[perm])
    ev_s = np.ascontiguousarray(ev4[perm])
    rho, pr = density_f64(p, pos_s)
    ev_f = ev_s.copy()
    ev_f[:, 3] = pr.astype(np.float32)
    pos_f = pos_s.copy()
    pos_f[:, 3] = (1.0 / rho).astype(np.float32)
    acc, scale, acc_alt = force_f64(p, pos_f, ev_s, rho, pr, alt=True)
    po, vo, eo = integrate_f64(p, pos_s, vel_s, ev_s, acc, rho)
    eo[:, 3] = pr
    # particles on the surface-force threshold: the other branch's outcome
    branch = np.any(acc_alt != acc, axis=1)
    pa, va = po, vo
    if branch.any():
        pa, va, _ = integrate_f64(p, pos_s, vel_s, ev_s, acc_alt, rho)
    return dict(grid=g, pos_sorted=pos_s, vel_sorted=vel_s, ev_sorted=ev_s, rho=rho,
                pressure=pr, acc=acc, acc_scale=scale, pos=po, vel=vo, ev=eo,
                branch=branch, acc_alt=acc_alt, pos_alt=pa, vel_alt=va)
