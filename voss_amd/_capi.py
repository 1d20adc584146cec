"""ctypes binding of libvoss.so (the MI355X build of the voss C ABI).

Loader order mirrors bindings/python/voss/_capi.py:14-46:
  1. $VOSS_LIBRARY_PATH (must exist),
  2. libvoss.so next to this package (the in-tree build),
  3. <repo>/library/build/libvoss.so (reference development layout).
The library is required: importing without it raises VossError, there is no
CPU stand-in.
"""

import ctypes
import os
from pathlib import Path

from .exceptions import VossError

VOSS_OK = 0
VOSS_ERR_INVALID_N = 1
VOSS_ERR_OUT_OF_RANGE = 2
VOSS_ERR_NO_CUDA = 3
VOSS_ERR_OUT_OF_MEMORY = 4
VOSS_ERR_CUDA = 5
VOSS_ERR_INVALID_ARG = 6
VOSS_ERR_INTERNAL = 7

VOSS_PROFILE_MINIMAL = 0
VOSS_PROFILE_STANDARD = 1
VOSS_PROFILE_FULL = 2


def _find_library() -> str:
    env = os.environ.get("VOSS_LIBRARY_PATH")
    if env:
        if not os.path.exists(env):
            raise VossError(f"VOSS_LIBRARY_PATH does not exist: {env}")
        return env
    here = Path(__file__).resolve().parent
    for cand in (here / "libvoss.so", here.parent / "library" / "build" / "libvoss.so"):
        if cand.exists():
            return str(cand)
    raise VossError("libvoss.so not found: build it with `python voss_amd/build.py` "
                    "or set VOSS_LIBRARY_PATH")


class VossPrimesStats(ctypes.Structure):
    _fields_ = [("mean_gap", ctypes.c_double), ("std_dev", ctypes.c_double),
                ("skewness", ctypes.c_double), ("kurtosis", ctypes.c_double),
                ("total_gaps", ctypes.c_uint64)]


class AmdGap(ctypes.Structure):
    _fields_ = [("position", ctypes.c_uint64), ("gap", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class AmdFullStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "prime_count", "pi_4_1", "pi_4_3", "pi_3_1", "pi_3_2", "last_prime", "max_gap",
        "n_large", "n_records", "n_checkpoints")]


class AmdRange(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("first", ctypes.c_uint64), ("last", ctypes.c_uint64),
                ("hist", ctypes.c_uint32 * 257), ("pad", ctypes.c_uint32)]


class AmdRequest(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("n", ctypes.c_uint32), ("a", ctypes.c_uint64),
                ("b", ctypes.c_uint64), ("x", ctypes.c_uint64 * 64)]


class AmdResponse(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("pad", ctypes.c_uint32), ("value", ctypes.c_uint64),
                ("first", ctypes.c_uint64), ("last", ctypes.c_uint64), ("hist", ctypes.c_uint32 * 257),
                ("pad2", ctypes.c_uint32), ("latency_ms", ctypes.c_double), ("tick", ctypes.c_uint64)]


class AmdServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int32), ("max_batch", ctypes.c_uint32), ("hbm_reserve", ctypes.c_uint64),
                ("hbm_budget", ctypes.c_uint64), ("nth_max", ctypes.c_uint64), ("goldbach_max", ctypes.c_uint64),
                ("plain", ctypes.c_int32), ("pad", ctypes.c_int32)]


class AmdServeStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("wall_s", "qps", "p50_ms", "p99_ms", "mean_ms", "max_ms",
                                               "warmup_s")] + \
               [(n, ctypes.c_uint64) for n in ("ticks", "max_tick", "deferred_by_memory", "peak_inflight_bytes")]


PROGRESS_CB = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                               ctypes.c_uint64, ctypes.c_void_p)

_u64, _u32, _int, _vp, _cp = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p
_P = ctypes.POINTER
_pu64, _pu32, _pint, _pu8 = _P(_u64), _P(_u32), _P(_int), _P(ctypes.c_uint8)
_ppu64 = _P(_P(_u64))

# symbol -> (argtypes, restype); every symbol is bound at import, so a library
# missing one fails here (as the reference binding does).
_SIGNATURES = {
    "voss_strerror": ([_int], _cp),
    "voss_get_last_error": ([], _cp),
    "voss_free": ([_vp], None),
    "voss_primes_prime_count": ([_u64, _pu64], _int),
    "voss_primes_in_range": ([_u64, _u64, _ppu64, _pu64], _int),
    "voss_primes_in_range_with_limit": ([_u64, _u64, _u64, _ppu64, _pu64], _int),
    "voss_primes_nth": ([_u64, _pu64], _int),
    "voss_primes_next": ([_u64, _pu64], _int),
    "voss_primes_prev": ([_u64, _pu64], _int),
    "voss_primes_is_prime": ([_u64, _pint], _int),
    "voss_primes_factorize": ([_u64, _ppu64, _pu64], _int),
    "voss_primes_is_mersenne_prime": ([_u32, _pint], _int),
    "voss_primes_is_fermat_prime": ([_u32, _pint], _int),
    "voss_primes_goldbach_count": ([_u64, _pu64], _int),
    "voss_primes_goldbach_partitions": ([_u64, _ppu64, _pu64], _int),
    "voss_primes_goldbach_test_upto": ([_u64, _ppu64, _pu64, _u64], _int),
    "voss_amd_goldbach_test_upto_pmax": ([_u64, _u64, _u64, _u64, _ppu64, _pu64, _pu64], _int),
    "voss_primes_ctx_new": ([_u64, _int, _P(_vp)], _int),
    "voss_primes_ctx_free": ([_vp], None),
    "voss_primes_ctx_prime_count": ([_vp, _pu64], _int),
    "voss_primes_ctx_twins": ([_vp, _pu64], _int),
    "voss_primes_ctx_cousin": ([_vp, _pu64], _int),
    "voss_primes_ctx_sexy": ([_vp, _pu64], _int),
    "voss_primes_ctx_statistics": ([_vp, _P(VossPrimesStats)], _int),
    "voss_primes_ctx_chebyshev_bias": ([_vp, _pu64, _pu64], _int),
    "voss_primes_ctx_large_gaps_count": ([_vp, _pu64], _int),
    "voss_primes_ctx_large_gaps_get": ([_vp, _u64, _pu64, _pu32], _int),
    "voss_primes_ctx_export_csv": ([_vp, _cp], _int),
    "voss_primes_ctx_set_progress": ([_vp, PROGRESS_CB, _vp], _int),
    "voss_primes_ctx_sophie_germain": ([_vp, _pu64], _int),
    "voss_primes_ctx_sophie_germain_upto": ([_vp, _pu64], _int),
    "voss_numbers_phi": ([_u64, _pu64], _int),
    "voss_numbers_tau": ([_u64, _pu64], _int),
    "voss_numbers_sigma": ([_u64, _pu64], _int),
    "voss_numbers_mu": ([_u64, _pint], _int),
    "voss_numbers_gcd": ([_u64, _u64, _pu64], _int),
    "voss_numbers_lcm": ([_u64, _u64, _pu64], _int),
    "voss_numbers_fibonacci": ([_u64, _pu64], _int),
    "voss_numbers_factorial": ([_u64, _pu64], _int),
    # extensions (include/voss/voss_amd.h)
    "voss_amd_full_stats": ([_u64, _u64, _u32, _P(AmdFullStats), _pu64, _pu64, _u64,
                             _P(AmdGap), _u64, _P(AmdGap), _u64], _int),
    "voss_amd_range_stats": ([_u64, _u64, _P(AmdRange)], _int),
    "voss_amd_batch_ranges": ([_pu64, _pu64, _u64, _P(AmdRange)], _int),
    "voss_amd_count_range": ([_u64, _u64, _pu64], _int),
    "voss_amd_count_range_noalloc": ([_u64, _u64, ctypes.c_void_p, ctypes.c_void_p], _int),
    "voss_amd_lds_probe": ([_int, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_double)], _int),
    "voss_amd_is_prime_batch": ([_pu64, _u64, _pu8], _int),
    "voss_amd_next_prime_batch": ([_pu64, _u64, _pu64], _int),
    "voss_amd_prev_prime_batch": ([_pu64, _u64, _pu64], _int),
    "voss_amd_nth_prime": ([_u64, _pu64], _int),
    "voss_amd_goldbach_counts": ([_pu64, _u64, _pu64], _int),
    "voss_amd_set_config": ([_u32, _int, _int, _u32], _int),
    "voss_amd_get_config": ([_pu32, _pint, _pint, _pu32], _int),
    "voss_amd_set_config_for": ([_int, _u32, _int, _int, _u32], _int),
    "voss_amd_get_config_for": ([_int, _pu32, _pint, _pint, _pu32], _int),
    "voss_amd_last_kernel_ms": ([_P(ctypes.c_double), _P(ctypes.c_double), _P(ctypes.c_double), _pu64], _int),
    "voss_amd_device_info": ([_pint, ctypes.c_char_p, _int], _int),
    "voss_amd_synchronize": ([], _int),
    "voss_amd_server_default_config": ([_P(AmdServerConfig)], _int),
    "voss_amd_server_start": ([_P(AmdServerConfig)], _int),
    "voss_amd_server_stop": ([], _int),
    "voss_amd_server_submit": ([_P(AmdRequest), _pu64], _int),
    "voss_amd_server_wait": ([_u64, _P(AmdResponse)], _int),
    "voss_amd_server_replay": ([_vp, _vp, _u64, _vp, _P(AmdServeStats)], _int),
}

_lib = ctypes.CDLL(_find_library())
for _name, (_args, _res) in _SIGNATURES.items():
    _fn = getattr(_lib, _name)
    _fn.argtypes = _args
    _fn.restype = _res
LIBRARY_PATH = _lib._name


def _check(rc: int, errcode_map: dict) -> None:
    """Raise the mapped exception for a nonzero code: "<strerror>: <last error>"."""
    if rc == VOSS_OK:
        return
    cls = errcode_map.get(rc, VossError)
    detail = _lib.voss_get_last_error().decode("utf-8", errors="replace")
    short = _lib.voss_strerror(rc).decode("utf-8", errors="replace")
    raise cls(f"{short}: {detail}")


def u64_array(values):
    arr = (ctypes.c_uint64 * max(1, len(values)))()
    arr[:len(values)] = list(values)
    return arr


def take_u64(arr, n):
    """Copy a library-owned uint64 array into a list and release it."""
    if n == 0 or not arr:
        return []
    try:
        return [int(arr[i]) for i in range(n)]
    finally:
        _lib.voss_free(arr)
