"""hmd_nbl_pair_snapshot_noalloc (the pair snapshot staged in a fixed device
scratch array, for callers that cannot allocate) against
hmd_nbl_pair_snapshot, through the C ABI: the same pair sets on two calls in
a row, the second on a smaller system, and a cap above the scratch maximum
refused with HMD_ERR_INVALID before any device work."""
import ctypes

import numpy as np
import pytest

from heddlemd_amd._lib import HmdError, check, lib
from tests.gpu_helpers import water_case

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 21
HMD_ERR_INVALID = 2


def _fn():
    f = lib().hmd_nbl_pair_snapshot_noalloc
    f.restype = ctypes.c_int32
    f.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint64,
                  ctypes.POINTER(ctypes.c_uint64)]
    return f


def _sorted(pairs):
    key = np.sort((pairs[:, 0].astype(np.uint64) << np.uint64(32)) | pairs[:, 1].astype(np.uint64))
    return key


def _noalloc_snapshot(field, cap):
    out = np.zeros((max(cap, 1), 2), np.uint32)
    k = ctypes.c_uint64()
    check(_fn()(field.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap,
                ctypes.byref(k)))
    return out[:min(k.value, cap)], int(k.value)


def test_noalloc_snapshot_equals_the_allocating_one(gpu_ctx):
    for grid, relax in (((12, 12, 9), 40), ((8, 8, 6), 20)):
        sysw, ff, part, field, integ = water_case(gpu_ctx, grid=grid, relax=relax)
        try:
            field.request_rebuild()
            field.compute(0)
            want = field.pair_snapshot()
            got, total = _noalloc_snapshot(field, len(want))
            print(f"{sysw['n']} atoms: {total} pairs (allocating entry point {len(want)})")
            assert total == len(want)
            assert np.array_equal(_sorted(got), _sorted(want))
            # a cap below the total stores a prefix and still reports the total
            part_got, total2 = _noalloc_snapshot(field, 1000)
            assert total2 == total and len(part_got) == 1000
            assert np.isin(_sorted(part_got), _sorted(want)).all()
        finally:
            for h in (integ, field, part):
                h.close()


def test_noalloc_snapshot_refuses_caps_beyond_the_scratch(gpu_ctx):
    sysw, ff, part, field, integ = water_case(gpu_ctx, grid=(8, 8, 6))
    try:
        k = ctypes.c_uint64()
        buf = np.zeros((1, 2), np.uint32)
        with pytest.raises(HmdError) as e:
            check(_fn()(field.h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                        MAX_PAIRS + 1, ctypes.byref(k)))
        assert e.value.code == HMD_ERR_INVALID
        assert "exceeds the no-allocation scratch" in str(e.value)
    finally:
        for h in (integ, field, part):
            h.close()
