"""The no-allocation count (voss_amd_count_range_noalloc) against the entry
point it varies (voss_amd_count_range), through the C ABI: results land in a
caller-owned device buffer, on the null stream and on a caller's stream, for
consecutive calls whose second range is smaller (the scratch array keeps the
last call's contents), plus its argument and size checks."""

import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _hip():
    # the HIP runtime libvoss.so already loaded (never a second copy)
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert paths, "libamdhip64 is not loaded"
    h = ctypes.CDLL(sorted(paths)[0])
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return h


H2D, D2H = 1, 2


def _read(hip, d):
    v = ctypes.c_uint64()
    assert hip.hipMemcpy(ctypes.byref(v), d, 8, D2H) == 0
    return v.value


def _write(hip, d, x):
    v = ctypes.c_uint64(x)
    assert hip.hipMemcpy(d, ctypes.byref(v), 8, H2D) == 0


def test_noalloc_count_equals_count_range(voss, oracle):
    from voss_amd import _capi
    L = _capi._lib
    voss.stats.device_info()  # initialises the library on the device (allocations allowed here)
    hip = _hip()
    d = ctypes.c_void_p()
    st = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), 64) == 0
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    try:
        # each pair: a wide call, then a smaller one
        cases = [(10**12, 10**12 + 5 * 10**10), (2, 10**9), (10**13, 10**13 + 10**6), (1, 10),
                 (0, 1), (7, 7), (1_000_000_007, 1_000_000_007), (10**14 - 10**9, 10**14)]
        for i, (a, b) in enumerate(cases):
            want = ctypes.c_uint64()
            assert L.voss_amd_count_range(a, b, ctypes.byref(want)) == 0
            stream = st if i % 2 else None
            _write(hip, d, 0xDEADBEEF)
            assert L.voss_amd_count_range_noalloc(a, b, d, stream) == 0, L.voss_get_last_error()
            assert (hip.hipStreamSynchronize(stream) if stream else hip.hipDeviceSynchronize()) == 0
            assert _read(hip, d) == want.value, (a, b)
        # the existing entry point itself against the oracle on two of them
        assert _read(hip, d) == oracle.range_stats(10**14 - 10**9, 10**14)["count"]
        assert L.voss_amd_count_range_noalloc(2, 10**9, d, None) == 0
        assert hip.hipDeviceSynchronize() == 0
        assert _read(hip, d) == 50_847_534
        # too many segments for the scratch array: a clear error, d_out untouched
        _write(hip, d, 12345)
        assert L.voss_amd_count_range_noalloc(2, 10**12, d, None) == _capi.VOSS_ERR_OUT_OF_RANGE
        assert b"at most 16384" in L.voss_get_last_error()
        assert L.voss_amd_count_range_noalloc(0, 10**14 + 1, d, None) == _capi.VOSS_ERR_OUT_OF_RANGE
        assert L.voss_amd_count_range_noalloc(0, 100, None, None) == _capi.VOSS_ERR_INVALID_ARG
        assert hip.hipDeviceSynchronize() == 0
        assert _read(hip, d) == 12345
    finally:
        hip.hipStreamDestroy(st)
        hip.hipFree(d)
