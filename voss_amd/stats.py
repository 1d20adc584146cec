"""Full-statistics sweep and batched range queries (no reference counterpart:
these expose the one-pass outputs of the MI355X engine; see
include/voss/voss_amd.h)."""

import ctypes

from . import _capi
from .primes import _ERRMAP, VossInvalidNError

_L = _capi._lib


def full_stats(N: int, cp_interval: int = 0, large_threshold: int = 500,
               with_hist: bool = True, rec_cap: int = 4096, large_cap: int = 4096,
               as_numpy: bool = False) -> dict:
    """Every statistic of [2, N] from one GPU sweep.

    Returns a dict: prime_count, pi_4_1, pi_4_3, pi_3_1, pi_3_2, last_prime,
    max_gap, hist (list, index = gap; hist[1] counts 2->3), checkpoints (one
    tuple (pi, pi_4_1, pi_4_3, pi_3_1, pi_3_2) per x = j * cp_interval),
    large_gaps ((position, gap) with gap >= large_threshold, largest first, at
    most large_cap), records ((position, gap) maximal gaps ascending), n_large,
    n_records.  as_numpy returns hist as a numpy uint64 array and checkpoints
    as an (n, 5) numpy uint64 array (views, no per-element conversion).
    """
    if not isinstance(N, int):
        raise TypeError("N must be int")
    if N < 2:
        raise VossInvalidNError(f"N must be >= 2, got {N}")
    st = _capi.AmdFullStats()
    if with_hist and as_numpy:
        # uninitialised: the library writes all 100000 entries
        import numpy as np
        hist_np = np.empty(100000, dtype=np.uint64)
        hist = hist_np.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    else:
        hist = (ctypes.c_uint64 * 100000)() if with_hist else None
    ncp = N // cp_interval if cp_interval else 0
    cps = (ctypes.c_uint64 * max(1, 5 * ncp))()
    # the gap lists are read only up to the counts the library returns, so
    # their buffers need no zero fill
    import numpy as np
    lg_buf = np.empty(max(1, large_cap) * ctypes.sizeof(_capi.AmdGap), dtype=np.uint8)
    rec_buf = np.empty(max(1, rec_cap) * ctypes.sizeof(_capi.AmdGap), dtype=np.uint8)
    lg = lg_buf.ctypes.data_as(ctypes.POINTER(_capi.AmdGap))
    rec = rec_buf.ctypes.data_as(ctypes.POINTER(_capi.AmdGap))
    rc = _L.voss_amd_full_stats(N, cp_interval, large_threshold, ctypes.byref(st), hist,
                                cps if ncp else None, ncp, lg, large_cap, rec, rec_cap)
    _capi._check(rc, _ERRMAP)
    nl, nr = min(st.n_large, large_cap), min(st.n_records, rec_cap)
    return {
        "N": N,
        "prime_count": st.prime_count, "pi_4_1": st.pi_4_1, "pi_4_3": st.pi_4_3,
        "pi_3_1": st.pi_3_1, "pi_3_2": st.pi_3_2, "last_prime": st.last_prime,
        "max_gap": st.max_gap, "n_large": st.n_large, "n_records": st.n_records,
        "hist": (hist_np if as_numpy else list(hist)) if with_hist else None,
        "checkpoints": (_np_view(cps)[:5 * ncp].reshape(ncp, 5) if as_numpy
                        else [tuple(cps[5 * j + f] for f in range(5)) for j in range(ncp)]),
        "large_gaps": [(lg[i].position, lg[i].gap) for i in range(nl)],
        "records": [(rec[i].position, rec[i].gap) for i in range(nr)],
    }


def _np_view(arr):
    import numpy as np
    return np.frombuffer(arr, dtype=np.uint64)


def _range_dict(r):
    return {"count": r.count, "first": r.first, "last": r.last, "hist": list(r.hist)}


def range_stats(a: int, b: int) -> dict:
    """Count, first, last and gap histogram (bins 2..512; [0] = other gaps) of
    the primes in [a, b]: one query, one launch."""
    out = _capi.AmdRange()
    _capi._check(_L.voss_amd_range_stats(a, b, ctypes.byref(out)), _ERRMAP)
    return _range_dict(out)


_RANGE_DTYPE = None


def _range_dtype():
    global _RANGE_DTYPE
    if _RANGE_DTYPE is None:
        import numpy as np
        _RANGE_DTYPE = np.dtype([("count", "<u8"), ("first", "<u8"), ("last", "<u8"),
                                 ("hist", "<u4", (257,)), ("pad", "<u4")])
    return _RANGE_DTYPE


def _u64_arg(v):
    try:
        import numpy as np
        if isinstance(v, np.ndarray):
            arr = np.ascontiguousarray(v, dtype=np.uint64)
            return arr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    except ImportError:
        pass
    arr = _capi.u64_array(v)
    return arr, arr


def batch_ranges(a, b, as_numpy: bool = False):
    """range_stats for many queries in one persistent launch.

    as_numpy returns a numpy structured array (fields count, first, last,
    hist[257]) viewing the results without per-element conversion."""
    n = len(a)
    if len(b) != n:
        raise ValueError("a and b must have the same length")
    ka, pa = _u64_arg(a)
    kb, pb = _u64_arg(b)
    if as_numpy:
        # the library writes every field of every record: no zero fill
        import numpy as np
        arr = np.empty(max(1, n), dtype=_range_dtype())
        _capi._check(_L.voss_amd_batch_ranges(pa, pb, n, arr.ctypes.data_as(ctypes.POINTER(_capi.AmdRange))),
                     _ERRMAP)
        return arr[:n]
    out = (_capi.AmdRange * max(1, n))()
    _capi._check(_L.voss_amd_batch_ranges(pa, pb, n, out), _ERRMAP)
    return [_range_dict(out[i]) for i in range(n)]


def count_range(a: int, b: int) -> int:
    out = ctypes.c_uint64(0)
    _capi._check(_L.voss_amd_count_range(a, b, ctypes.byref(out)), _ERRMAP)
    return out.value


def is_prime_batch(xs: list) -> list:
    """GPU Miller-Rabin over many u64 values."""
    n = len(xs)
    out = (ctypes.c_uint8 * max(1, n))()
    _capi._check(_L.voss_amd_is_prime_batch(_capi.u64_array(xs), n, out), _ERRMAP)
    return [bool(out[i]) for i in range(n)]


def next_prime_batch(xs, as_numpy: bool = False):
    """Smallest prime > x per value (GPU; 0 where none fits in uint64).
    as_numpy returns a numpy uint64 array (xs may be one too)."""
    n = len(xs)
    out = (ctypes.c_uint64 * max(1, n))()
    keep, px = _u64_arg(xs)
    _capi._check(_L.voss_amd_next_prime_batch(px, n, out), _ERRMAP)
    if as_numpy:
        import numpy as np
        return np.frombuffer(out, dtype=np.uint64, count=n).copy()
    return list(out[:n])


def prev_prime_batch(xs: list) -> list:
    """Largest prime < x per value (0 for x <= 2)."""
    n = len(xs)
    out = (ctypes.c_uint64 * max(1, n))()
    _capi._check(_L.voss_amd_prev_prime_batch(_capi.u64_array(xs), n, out), _ERRMAP)
    return list(out[:n])


def nth_prime(n: int) -> int:
    """n-th prime without the 5e7 cap of primes.nth (n <= 1e12)."""
    out = ctypes.c_uint64(0)
    _capi._check(_L.voss_amd_nth_prime(n, ctypes.byref(out)), _ERRMAP)
    return out.value


def goldbach_counts(ns: list) -> list:
    """Goldbach partition counts for many even n in one launch."""
    k = len(ns)
    out = (ctypes.c_uint64 * max(1, k))()
    _capi._check(_L.voss_amd_goldbach_counts(_capi.u64_array(ns), k, out), _ERRMAP)
    return list(out[:k])


KIND = {"both": -1, "sweep": 0, "batch": 1}


def goldbach_test_upto(N: int, max_ce: int = 10, pmax: int = 0, scan_limit: int = 0):
    """(total, first max_ce counterexamples ascending) among the even n <= N
    when Goldbach partners are restricted to primes <= pmax (0: none); the
    first pass tries partners up to scan_limit (0: all sieving primes) and
    settles the rest by exact counts."""
    arr = ctypes.POINTER(ctypes.c_uint64)()
    cnt, tot = ctypes.c_uint64(0), ctypes.c_uint64(0)
    _capi._check(_L.voss_amd_goldbach_test_upto_pmax(N, pmax, scan_limit, max_ce, ctypes.byref(arr),
                                                     ctypes.byref(cnt), ctypes.byref(tot)), _ERRMAP)
    return tot.value, _capi.take_u64(arr, cnt.value)


def set_config(seg_bytes: int, block: int = 1024, wgs_per_cu: int = 1, tier1_p: int = 2048,
               kind: str = "both") -> None:
    """Engine launch configuration for sweeps, batched ranges, or both."""
    _capi._check(_L.voss_amd_set_config_for(KIND[kind], seg_bytes, block, wgs_per_cu, tier1_p), _ERRMAP)


def get_config(kind: str = "sweep") -> dict:
    s, b, w, t = ctypes.c_uint32(), ctypes.c_int(), ctypes.c_int(), ctypes.c_uint32()
    _capi._check(_L.voss_amd_get_config_for(KIND[kind], ctypes.byref(s), ctypes.byref(b), ctypes.byref(w),
                                            ctypes.byref(t)), _ERRMAP)
    return {"seg_bytes": s.value, "block": b.value, "wgs_per_cu": w.value, "tier1_p": t.value}


def last_kernel_ms() -> dict:
    a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    s = ctypes.c_uint64()
    _capi._check(_L.voss_amd_last_kernel_ms(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                            ctypes.byref(s)), _ERRMAP)
    return {"sieve_ms": a.value, "stitch_ms": b.value, "total_ms": c.value, "segments": s.value}


def lds_probe(kind: int, iters: int = 20000) -> dict:
    """LDS lane-operations per clock per CU for one access kind (0: ds_and
    conflict-free, 1: ds_and random, 2: ds_read conflict-free, 3: ds_read
    random), the shader clock measured meanwhile and the kernel time."""
    r, c, m = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _capi._check(_L.voss_amd_lds_probe(kind, iters, ctypes.byref(r), ctypes.byref(c), ctypes.byref(m)), _ERRMAP)
    return {"ops_per_clk_cu": r.value, "clock_hz": c.value, "ms": m.value}


def device_info() -> dict:
    n = ctypes.c_int()
    name = ctypes.create_string_buffer(64)
    _capi._check(_L.voss_amd_device_info(ctypes.byref(n), name, 64), _ERRMAP)
    return {"num_cus": n.value, "arch": name.value.decode()}


def synchronize() -> None:
    _capi._check(_L.voss_amd_synchronize(), _ERRMAP)
