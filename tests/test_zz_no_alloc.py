"""The no-allocation entry point gsph_checksum_ranges_noalloc against the
allocating gsph_checksum_range: bit-identical per range on two calls in a
row (the second with fewer, smaller ranges), results written to a device
buffer the test allocates, and no device memory taken by the call."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    lib.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


def test_checksum_ranges_noalloc_matches_checksum_range(gsph):
    h = hip()
    params, _ = gsph.scene_config(0)
    worlds, blocks = gsph.batch_workload(12, seed=31)
    counts = [len(b) for b in blocks]
    print(f"\n[gpu-case] no-alloc checksum scenes=12 n={sum(counts)}", flush=True)
    b = gsph.System(params, np.concatenate(blocks), batch_worlds=worlds, batch_counts=counts)
    b.step(2)
    b.sync()
    ranges = [b.scene_range(s) for s in range(12)]
    buf = C.c_void_p()
    assert h.hipMalloc(C.byref(buf), 8 * 5 * 64) == 0
    try:
        calls = [
            ([r[0] for r in ranges], [r[1] for r in ranges]),
            # second call: fewer ranges, each a smaller part of a scene
            ([ranges[k][0] + 7 for k in (1, 4, 9)], [ranges[k][1] // 3 for k in (1, 4, 9)]),
        ]
        for firsts, cnts in calls:
            assert h.hipMemset(buf, 0xff, 8 * 5 * 64) == 0
            assert h.hipDeviceSynchronize() == 0
            free0, total = C.c_size_t(), C.c_size_t()
            assert h.hipMemGetInfo(C.byref(free0), C.byref(total)) == 0
            b.checksum_ranges_noalloc(firsts, cnts, buf.value)
            b.sync()
            free1 = C.c_size_t()
            assert h.hipMemGetInfo(C.byref(free1), C.byref(total)) == 0
            assert free1.value == free0.value          # the call took no device memory
            got = np.zeros(5 * len(firsts), np.float64)
            assert h.hipMemcpy(got.ctypes.data_as(C.c_void_p), buf, got.nbytes, 2) == 0  # D2H
            got = got.reshape(-1, 5)
            for k, (f, c) in enumerate(zip(firsts, cnts)):
                exp = b.checksum(f, c)
                assert np.array_equal(got[k], exp), (k, got[k], exp)
        with pytest.raises(gsph.GsphError):
            b.checksum_ranges_noalloc([0] * 65, [1] * 65, buf.value)   # beyond the scratch
    finally:
        h.hipFree(buf)
        b.close()
