"""ctypes binding of libgsph.so (the C ABI declared in include/gsph.h).

Host arrays are numpy arrays: pos4 (n, 4) float32 with w = 1/density,
vel3 (n, 3) float32, ev4 (n, 4) float32 with w = pressure, acc3 (n, 3).
Every call raises GsphError on a non-zero status; nothing here falls back to
a CPU path - without the built library the import fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("GSPH_LIBRARY", PKG_DIR / "_build" / "libgsph.so"))

ERRORS = {
    -1: "GSPH_ERR_INVALID_ARG",
    -2: "GSPH_ERR_HIP",
    -3: "GSPH_ERR_OUT_OF_DOMAIN",
    -4: "GSPH_ERR_CAPACITY",
    -5: "GSPH_ERR_NO_DEVICE",
    -6: "GSPH_ERR_SPIN_TIMEOUT",
    -7: "GSPH_ERR_IO",
    -8: "GSPH_ERR_BUDGET",
    -9: "GSPH_ERR_INJECTED",
}


class GsphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("kernel_h", C.c_float), ("mass", C.c_float), ("rest_density", C.c_float),
        ("gas_constant", C.c_float), ("viscosity", C.c_float), ("time_step", C.c_float),
        ("gravity", C.c_float * 3), ("wall_damping", C.c_float),
        ("surface_normal", C.c_float), ("surface_coe", C.c_float),
        ("world_size", C.c_float * 3),
    ]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else float(v)
        return d


class Constants(C.Structure):
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


class Options(C.Structure):
    _fields_ = [
        ("split_mode", C.c_int), ("sms_min_self", C.c_int), ("sms_min_total", C.c_int),
        ("blocks_per_cu", C.c_int), ("pair_masks", C.c_int), ("row_prune", C.c_int),
        ("stage_depth", C.c_int), ("sms_mask_bytes", C.c_int), ("sms_local_merge", C.c_int),
        ("sms_legacy_staging", C.c_int), ("same_cell_pair_force", C.c_int),
        ("lean_passes", C.c_int),
    ]


class StageMs(C.Structure):
    _fields_ = [("arrange", C.c_float), ("density", C.c_float), ("force", C.c_float),
                ("integrate", C.c_float), ("total", C.c_float)]


class PairStats(C.Structure):
    _fields_ = [(k, C.c_int) for k in (
        "total_pairs", "same_cell_pairs", "is_same_pairs", "full_is_same_pairs",
        "partial_is_same_pairs", "compact_is_same_pairs", "widened_is_same_pairs")]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_int), ("failed", C.c_int), ("rejected", C.c_int),
                ("max_active", C.c_int), ("p50", C.c_double), ("p99", C.c_double),
                ("wall", C.c_double), ("particle_steps", C.c_double),
                ("engine_steps", C.c_longlong), ("recompositions", C.c_longlong),
                ("peak_bytes", C.c_double)]


REQ_STATES = {0: "queued", 1: "active", 2: "done", 3: "failed", 4: "rejected"}

# 32-byte task record (gsph_task / the reference's BlockTask)
TASK_DTYPE = np.dtype([
    ("is_same", "i1"), ("pad0", "u1"), ("paired_bounds", "<u2"), ("cellid", "<i4"),
    ("cell_pos", "<u2", (3,)), ("p_offset", "<u2"),
    ("xxi", "i1"), ("xxx", "i1"), ("yyi", "i1"), ("yyy", "i1"), ("zzi", "i1"), ("zzz", "i1"),
    ("pad1", "i1", (2,)), ("cell_begin", "<i4"), ("cell_particle_count", "<i4"),
])
assert TASK_DTYPE.itemsize == 32

_P = C.c_void_p
_I = C.c_int
_SIGS = {
    "gsph_version": (C.c_char_p, []),
    "gsph_device_count": (_I, [C.POINTER(_I)]),
    "gsph_last_error": (C.c_char_p, [_P]),
    "gsph_params_default": (None, [C.POINTER(Params)]),
    "gsph_options_default": (None, [C.POINTER(Options)]),
    "gsph_derive_constants": (_I, [C.POINTER(Params), C.POINTER(Constants)]),
    "gsph_scene_json": (_I, [C.c_char_p, C.POINTER(Params), C.POINTER(_I), _P, _I]),
    "gsph_scene_config": (_I, [_I, C.POINTER(Params), C.POINTER(_I), _P, _I]),
    "gsph_scene_block": (_I, [_P, _I, _I, _I, C.c_float, C.c_float, C.c_uint64, _P]),
    "gsph_batch_create": (_I, [C.POINTER(Params), C.POINTER(Options), _I, _P, _P, _P, _P,
                               C.POINTER(_P)]),
    "gsph_scene_count": (_I, [_P, C.POINTER(_I)]),
    "gsph_scene_range": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I)]),
    "gsph_scene_world": (_I, [_P, _I, C.POINTER(C.c_float)]),
    "gsph_get_params": (_I, [_P, C.POINTER(Params)]),
    "gsph_is_batched": (_I, [_P, C.POINTER(_I)]),
    "gsph_step_count": (_I, [_P, C.POINTER(C.c_longlong)]),
    "gsph_set_step_count": (_I, [_P, C.c_longlong]),
    "gsph_snapshot_save": (_I, [_P, C.c_char_p]),
    "gsph_snapshot_check": (_I, [C.c_char_p, C.POINTER(C.c_longlong), C.POINTER(_I),
                                 C.POINTER(C.c_longlong)]),
    "gsph_snapshot_load": (_I, [C.c_char_p, C.POINTER(Options), C.POINTER(_P)]),
    "gsph_create": (_I, [C.POINTER(Params), C.POINTER(Options), _I, _P, _P, C.POINTER(_P)]),
    "gsph_destroy": (_I, [_P]),
    "gsph_num_particles": (_I, [_P, C.POINTER(_I)]),
    "gsph_get_constants": (_I, [_P, C.POINTER(Constants)]),
    "gsph_set_state": (_I, [_P, _P, _P, _P]),
    "gsph_add_particles": (_I, [_P, _I, _P, _P, _I]),
    "gsph_arrange": (_I, [_P]),
    "gsph_density": (_I, [_P]),
    "gsph_force": (_I, [_P]),
    "gsph_integrate": (_I, [_P]),
    "gsph_step": (_I, [_P, _I, _P]),
    "gsph_enable_graph": (_I, [_P, _I]),
    "gsph_autotune": (_I, [_P, _I, _I, C.POINTER(Options), C.POINTER(C.c_float)]),
    "gsph_set_options": (_I, [_P, C.POINTER(Options)]),
    "gsph_get_options": (_I, [_P, C.POINTER(Options)]),
    "gsph_sync": (_I, [_P]),
    "gsph_read_state": (_I, [_P, _P, _P, _P, _P]),
    "gsph_grid_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "gsph_mask_pool_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "gsph_read_grid": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "gsph_checksum": (_I, [_P, _P]),
    "gsph_checksum_range": (_I, [_P, _I, _I, _P]),
    "gsph_checksum_ranges_noalloc": (_I, [_P, _I, _P, _P, _P]),
    "gsph_get_pair_stats": (_I, [_P, C.POINTER(PairStats)]),
    "gsph_kernel_timing": (_I, [_P, _I]),
    "gsph_kernel_times": (_I, [_P, C.c_char_p, _I, _P, _P, _I, C.POINTER(_I)]),
    "gsph_inject_fault": (_I, [_P, _I]),
    "gsph_engine_bytes": (_I, [C.POINTER(Options), C.c_longlong, C.c_longlong, _I,
                               C.POINTER(C.c_double)]),
    "gsph_server_create": (_I, [C.POINTER(Params), C.POINTER(Options), C.c_double, _I,
                                C.POINTER(_P)]),
    "gsph_server_destroy": (_I, [_P]),
    "gsph_server_submit": (_I, [_P, _P, _I, _P, _P, _I, C.c_double, C.POINTER(_I)]),
    "gsph_server_run": (_I, [_P, _I]),
    "gsph_server_result": (_I, [_P, _I, _P, _P, _P, C.POINTER(_I), C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]),
    "gsph_server_get_stats": (_I, [_P, C.POINTER(ServerStats)]),
    "gsph_server_last_error": (C.c_char_p, [_P]),
}

# Every symbol include/gsph.h declares (checked by the CPU tests).
EXPORTED = sorted(_SIGS)

_lib = None


def load(path: str | os.PathLike | None = None):
    """Load libgsph.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} is not built; run __graft_entry__.build()")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc, ctx=None):
    if rc != 0:
        msg = load().gsph_last_error(ctx)
        raise GsphError(rc, msg.decode() if msg else "")


def params_default() -> Params:
    p = Params()
    load().gsph_params_default(C.byref(p))
    return p


def options_default() -> Options:
    o = Options()
    load().gsph_options_default(C.byref(o))
    return o


def derive_constants(p: Params) -> Constants:
    c = Constants()
    _check(load().gsph_derive_constants(C.byref(p), C.byref(c)))
    return c


def device_count() -> int:
    n = C.c_int(0)
    _check(load().gsph_device_count(C.byref(n)))
    return n.value


PROBE_PATH = PKG_DIR / "_build" / "libgsph_probe.so"
_PROBE_SIGS = {
    "gsph_probe_last_error": (C.c_char_p, []),
    "gsph_probe_copy": (_I, [C.c_double, _I, C.POINTER(C.c_float)]),
    "gsph_probe_valu": (_I, [_I, C.POINTER(C.c_float), C.POINTER(C.c_double)]),
}
# Every symbol include/gsph_probe.h declares (checked by the CPU tests).
PROBE_EXPORTED = sorted(_PROBE_SIGS)
_probe = None


def load_probe():
    """The measurement-only library libgsph_probe.so (calibration kernels)."""
    global _probe
    if _probe is None:
        if not PROBE_PATH.exists():
            raise FileNotFoundError(f"{PROBE_PATH} is not built; run __graft_entry__.build()")
        lib = C.CDLL(str(PROBE_PATH))
        for name, (res, args) in _PROBE_SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _probe = lib
    return _probe


def _check_probe(rc):
    if rc != 0:
        msg = load_probe().gsph_probe_last_error()
        raise GsphError(rc, msg.decode() if msg else "")


def probe_copy(nbytes: float, reps: int = 10) -> float:
    """Average ms of a plain copy kernel over nbytes (HBM counter calibration)."""
    ms = C.c_float()
    _check_probe(load_probe().gsph_probe_copy(float(nbytes), int(reps), C.byref(ms)))
    return ms.value


def probe_valu(iters: int = 4000):
    """Peak FP32 VALU rate: (ms, lane operations) of one FMA-chain launch."""
    ms, ops = C.c_float(), C.c_double()
    _check_probe(load_probe().gsph_probe_valu(int(iters), C.byref(ms), C.byref(ops)))
    return ms.value, ops.value


def engine_bytes(particles: int, cells: int, scenes: int = 1, options: Options | None = None) -> float:
    """Device bytes of a context with these capacities (host only)."""
    b = C.c_double()
    _check(load().gsph_engine_bytes(None if options is None else C.byref(options), int(particles),
                                    int(cells), int(scenes), C.byref(b)))
    return b.value


def scene_config(k: int):
    """Synthetic workload k (0, 1, 2) -> (Params, pos4 array)."""
    lib = load()
    p = Params()
    n = C.c_int(0)
    _check(lib.gsph_scene_config(k, C.byref(p), C.byref(n), None, 0))
    pos = np.zeros((n.value, 4), np.float32)
    _check(lib.gsph_scene_config(k, C.byref(p), C.byref(n), _ptr(pos), n.value))
    return p, pos


def scene_json(path):
    lib = load()
    p = Params()
    n = C.c_int(0)
    _check(lib.gsph_scene_json(str(path).encode(), C.byref(p), C.byref(n), None, 0))
    pos = np.zeros((n.value, 4), np.float32)
    _check(lib.gsph_scene_json(str(path).encode(), C.byref(p), C.byref(n), _ptr(pos), n.value))
    return p, pos


def scene_block(corner, nx, ny, nz, spacing, jitter_frac, seed):
    pos = np.zeros((nx * ny * nz, 4), np.float32)
    cor = np.asarray(corner, np.float32)
    _check(load().gsph_scene_block(_ptr(cor), nx, ny, nz, spacing, jitter_frac, seed, _ptr(pos)))
    return pos


def batch_workload(count=256, seed=20261206):
    """The batched dam-break workload: `count` lattice blocks (spacing 0.01,
    jitter 1%, corner 0.06) with N log-uniform in [4096, 262144] and aspect
    ratios y/x, z/x uniform in [0.5, 2], each in a box 2x its extent.
    Returns (worlds, blocks)."""
    rng = np.random.default_rng(seed)
    worlds, blocks = [], []
    for k in range(count):
        n_target = float(np.exp(rng.uniform(np.log(4096), np.log(262144))))
        ay, az = rng.uniform(0.5, 2.0, 2)
        nx = max(4, int(round((n_target / (ay * az)) ** (1 / 3))))
        ny, nz = max(4, int(round(ay * nx))), max(4, int(round(az * nx)))
        blocks.append(scene_block([0.06, 0.06, 0.06], nx, ny, nz, 0.01, 0.01, seed + 7919 * (k + 1)))
        worlds.append([2 * nx * 0.01, 2 * ny * 0.01, 2 * nz * 0.01])
    return worlds, blocks


def serving_workload(count=96, seed=20261207):
    """The serving workload's requests: N log-uniform in [8192, 1048576] (a
    cube lattice block of side round(N^(1/3)), spacing 0.01, jitter 1%,
    corner 0.06, in a cubic box 2x its extent) and a step count uniform in
    50..500.  Returns (worlds, blocks, steps, rng); the caller draws the
    arrival times from rng (Poisson at its chosen rate)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(8192), np.log(1048576), count))
    steps = rng.integers(50, 501, count)
    worlds, blocks = [], []
    for k in range(count):
        side = max(8, int(round(ns[k] ** (1 / 3))))
        blocks.append(scene_block([0.06, 0.06, 0.06], side, side, side, 0.01, 0.01, seed + k))
        w = 2 * side * 0.01
        worlds.append([w, w, w])
    return worlds, blocks, [int(x) for x in steps], rng


class System:
    """One context: a plain scene (gsph_create) or a batch (gsph_batch_create)."""

    def __init__(self, params: Params, pos4, vel3=None, options: Options | None = None,
                 *, batch_worlds=None, batch_counts=None, _handle=None):
        lib = load()
        self._lib = lib
        if _handle is not None:             # adopt a context made by the library
            self.h = _handle
            n = C.c_int()
            self._c(lib.gsph_num_particles(self.h, C.byref(n)))
            self.n = n.value
            self.params = Params()
            self._c(lib.gsph_get_params(self.h, C.byref(self.params)))
            return
        pos4 = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
        vel = None if vel3 is None else np.ascontiguousarray(vel3, np.float32).reshape(-1, 3)
        opt = options if options is not None else options_default()
        h = C.c_void_p()
        if batch_counts is None:
            _check(lib.gsph_create(C.byref(params), C.byref(opt), pos4.shape[0], _ptr(pos4),
                                   _ptr(vel), C.byref(h)))
        else:
            worlds = np.ascontiguousarray(batch_worlds, np.float32).reshape(-1, 3)
            counts = np.ascontiguousarray(batch_counts, np.int32)
            _check(lib.gsph_batch_create(C.byref(params), C.byref(opt), counts.shape[0],
                                         _ptr(worlds), _ptr(counts), _ptr(pos4), _ptr(vel),
                                         C.byref(h)))
        self.h = h
        self.n = pos4.shape[0]
        self.params = params

    def save(self, path):
        """Checkpoint: parameters, options, scene layout, step count, state."""
        self._c(self._lib.gsph_snapshot_save(self.h, os.fsencode(path)))

    def step_count(self) -> int:
        v = C.c_longlong()
        self._c(self._lib.gsph_step_count(self.h, C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            self._lib.gsph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self.h)

    def constants(self) -> Constants:
        c = Constants()
        self._c(self._lib.gsph_get_constants(self.h, C.byref(c)))
        return c

    def set_state(self, pos4, vel3=None, ev4=None):
        pos4 = np.ascontiguousarray(pos4, np.float32)
        vel3 = None if vel3 is None else np.ascontiguousarray(vel3, np.float32)
        ev4 = None if ev4 is None else np.ascontiguousarray(ev4, np.float32)
        self._c(self._lib.gsph_set_state(self.h, _ptr(pos4), _ptr(vel3), _ptr(ev4)))

    def add_particles(self, pos4, vel3=None, reset_all=False):
        """Append particles between steps (plain contexts; w of pos4 is ignored,
        ev starts at 0); reset_all also zeroes ev of the particles already there."""
        pos4 = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
        vel3 = None if vel3 is None else np.ascontiguousarray(vel3, np.float32).reshape(-1, 3)
        if vel3 is not None and vel3.shape[0] != pos4.shape[0]:
            raise ValueError("vel3 and pos4 differ in length")
        self._c(self._lib.gsph_add_particles(self.h, pos4.shape[0], _ptr(pos4), _ptr(vel3),
                                             1 if reset_all else 0))
        self.n += pos4.shape[0]

    def arrange(self):
        self._c(self._lib.gsph_arrange(self.h))

    def density(self):
        self._c(self._lib.gsph_density(self.h))

    def force(self):
        self._c(self._lib.gsph_force(self.h))

    def integrate(self):
        self._c(self._lib.gsph_integrate(self.h))

    def step(self, n=1, timings=False):
        if timings:
            out = (StageMs * max(n, 1))()
            self._c(self._lib.gsph_step(self.h, n, C.cast(out, C.c_void_p)))
            return [dict(arrange=o.arrange, density=o.density, force=o.force,
                         integrate=o.integrate, total=o.total) for o in out[:n]]
        self._c(self._lib.gsph_step(self.h, n, None))
        return None

    def enable_graph(self, on=True):
        self._c(self._lib.gsph_enable_graph(self.h, 1 if on else 0))

    def autotune(self, trial_steps=3, tune_split=False):
        """Pick launch options by timing this scene (state is restored);
        returns (options, ms per step)."""
        o, ms = Options(), C.c_float()
        self._c(self._lib.gsph_autotune(self.h, trial_steps, 1 if tune_split else 0,
                                        C.byref(o), C.byref(ms)))
        return o, ms.value

    def set_options(self, opt: Options):
        self._c(self._lib.gsph_set_options(self.h, C.byref(opt)))

    def options(self) -> Options:
        o = Options()
        self._c(self._lib.gsph_get_options(self.h, C.byref(o)))
        return o

    def sync(self):
        self._c(self._lib.gsph_sync(self.h))

    def read_state(self):
        n = self.n
        pos = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 3), np.float32)
        ev = np.zeros((n, 4), np.float32)
        acc = np.zeros((n, 3), np.float32)
        self._c(self._lib.gsph_read_state(self.h, _ptr(pos), _ptr(vel), _ptr(ev), _ptr(acc)))
        return pos, vel, ev, acc

    def grid_info(self):
        cells, tasks, mid, tra = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._c(self._lib.gsph_grid_info(self.h, C.byref(cells), C.byref(tasks), C.byref(mid),
                                         C.byref(tra)))
        return dict(cells=cells.value, tasks=tasks.value, middle_raw=mid.value, tra=tra.value)

    def mask_pool_info(self):
        """(rows taken by the last density pass, capacity) of the mask pool."""
        used, cap = C.c_int(), C.c_int()
        self._c(self._lib.gsph_mask_pool_info(self.h, C.byref(used), C.byref(cap)))
        return used.value, cap.value

    def read_grid(self):
        info = self.grid_info()
        n, ncell, T = self.n, info["cells"], info["tasks"]
        hash_ = np.zeros(n, np.int32)
        perm = np.zeros(n, np.int32)
        begin = np.zeros(ncell, np.int32)
        count = np.zeros(ncell, np.int32)
        req = np.zeros(ncell, np.int32)
        compact = np.zeros(n, np.int32)
        tasks = np.zeros(max(T + (T & 1), 1), TASK_DTYPE)
        self._c(self._lib.gsph_read_grid(self.h, _ptr(hash_), _ptr(perm), _ptr(begin),
                                         _ptr(count), _ptr(req), _ptr(compact), _ptr(tasks)))
        out = dict(info)
        out["num_tasks"] = out.pop("tasks")
        out.update(hash=hash_, perm=perm, begin=begin, count=count, req=req, compact=compact,
                   tasks=tasks[:T + (T & 1)])
        return out

    def checksum(self, first=None, count=None):
        out = np.zeros(5, np.float64)
        if first is None:
            self._c(self._lib.gsph_checksum(self.h, _ptr(out)))
        else:
            self._c(self._lib.gsph_checksum_range(self.h, first, count, _ptr(out)))
        return out

    def checksum_ranges_noalloc(self, firsts, counts, out5_device_ptr: int):
        """gsph_checksum_ranges_noalloc: enqueues the checksums of the ranges
        into a caller-owned device buffer (5 doubles per range); nothing is
        allocated and nothing waits (read the buffer after sync())."""
        f = np.ascontiguousarray(firsts, np.int32)
        c = np.ascontiguousarray(counts, np.int32)
        self._c(self._lib.gsph_checksum_ranges_noalloc(self.h, len(f), _ptr(f), _ptr(c),
                                                       C.c_void_p(out5_device_ptr)))

    def pair_stats(self):
        s = PairStats()
        self._c(self._lib.gsph_get_pair_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in PairStats._fields_}

    def scene_range(self, s):
        a, b = C.c_int(), C.c_int()
        self._c(self._lib.gsph_scene_range(self.h, s, C.byref(a), C.byref(b)))
        return a.value, b.value

    def kernel_timing(self, on=True):
        self._c(self._lib.gsph_kernel_timing(self.h, 1 if on else 0))

    def kernel_times(self):
        cap = 64
        names = C.create_string_buffer(4096)
        ms = np.zeros(cap, np.float32)
        launches = np.zeros(cap, np.int32)
        cnt = C.c_int()
        self._c(self._lib.gsph_kernel_times(self.h, names, 4096, _ptr(ms), _ptr(launches), cap,
                                            C.byref(cnt)))
        keys = names.value.decode().split(";") if cnt.value else []
        return {k: (float(ms[i]), int(launches[i])) for i, k in enumerate(keys)}

    def inject_fault(self, after_calls=0):
        self._c(self._lib.gsph_inject_fault(self.h, after_calls))


class Server:
    """Continuous-batching request loop (gsph_server_*)."""

    def __init__(self, params: Params, hbm_budget_bytes: float, max_scenes: int = 1024,
                 options: Options | None = None):
        lib = load()
        self._lib = lib
        opt = options if options is not None else options_default()
        h = C.c_void_p()
        _check(lib.gsph_server_create(C.byref(params), C.byref(opt), float(hbm_budget_bytes),
                                      int(max_scenes), C.byref(h)))
        self.h = h
        self.sizes = {}

    def _c(self, rc):
        if rc != 0:
            msg = self._lib.gsph_server_last_error(self.h)
            raise GsphError(rc, msg.decode() if msg else "")

    def submit(self, world, pos4, steps, arrival=0.0, vel3=None) -> int:
        w = np.ascontiguousarray(world, np.float32)
        pos4 = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
        vel = None if vel3 is None else np.ascontiguousarray(vel3, np.float32)
        rid = C.c_int()
        self._c(self._lib.gsph_server_submit(self.h, _ptr(w), pos4.shape[0], _ptr(pos4),
                                             _ptr(vel), int(steps), float(arrival),
                                             C.byref(rid)))
        self.sizes[rid.value] = pos4.shape[0]
        return rid.value

    def run(self, realtime=False):
        self._c(self._lib.gsph_server_run(self.h, 1 if realtime else 0))

    def result(self, rid):
        n = self.sizes[rid]
        pos = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 3), np.float32)
        ev = np.zeros((n, 4), np.float32)
        st = C.c_int()
        lat = C.c_double()
        q = C.c_double()
        self._c(self._lib.gsph_server_result(self.h, rid, _ptr(pos), _ptr(vel), _ptr(ev),
                                             C.byref(st), C.byref(lat), C.byref(q)))
        return dict(state=REQ_STATES[st.value], latency=lat.value, queued=q.value,
                    pos=pos, vel=vel, ev=ev)

    def stats(self):
        st = ServerStats()
        self._c(self._lib.gsph_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            self._lib.gsph_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def snapshot_check(path):
    """Validates a snapshot file without the GPU: (n, scenes, steps)."""
    lib = load()
    n, ns, st = C.c_longlong(), C.c_int(), C.c_longlong()
    _check(lib.gsph_snapshot_check(os.fsencode(path), C.byref(n), C.byref(ns), C.byref(st)))
    return n.value, ns.value, st.value


def load_snapshot(path, options: Options | None = None) -> System:
    """A context restored from a snapshot (continues bit-exactly)."""
    lib = load()
    h = C.c_void_p()
    _check(lib.gsph_snapshot_load(os.fsencode(path), None if options is None else C.byref(options),
                                  C.byref(h)))
    return System(None, None, _handle=h)
