"""Prime counting, gap statistics and range queries on the MI355X engine.

Same public surface and error behaviour as the reference package
(bindings/python/voss/primes.py): type checks raise TypeError (bool counts
as int), N < 2 raises VossInvalidNError before any library call, unknown
profile strings raise ValueError, a closed Context raises VossError, and
Context results are cached so repeated calls return the same object.
"""

import ctypes

from . import _capi
from ._capi import (VOSS_ERR_INVALID_ARG, VOSS_ERR_INVALID_N, VOSS_PROFILE_FULL,
                    VOSS_PROFILE_MINIMAL, VOSS_PROFILE_STANDARD)
from .exceptions import VossError


class VossInvalidNError(VossError):
    """Raised when N (or n, x, p) is outside the function's domain."""


class VossInvalidArgError(VossError):
    """Raised when an argument is invalid (bad profile, MINIMAL query, ...)."""


_ERRMAP = {VOSS_ERR_INVALID_N: VossInvalidNError, VOSS_ERR_INVALID_ARG: VossInvalidArgError}
_PROFILES = {"minimal": VOSS_PROFILE_MINIMAL, "standard": VOSS_PROFILE_STANDARD,
             "full": VOSS_PROFILE_FULL}
_L = _capi._lib


def _need_int(name, v):
    if not isinstance(v, int):
        raise TypeError(f"{name} must be int, got {type(v).__name__}")


def _call_u64(fn, *args):
    out = ctypes.c_uint64(0)
    _capi._check(fn(*args, ctypes.byref(out)), _ERRMAP)
    return out.value


def _call_flag(fn, *args):
    out = ctypes.c_int(0)
    _capi._check(fn(*args, ctypes.byref(out)), _ERRMAP)
    return out.value == 1


def _call_array(fn, *args):
    arr = ctypes.POINTER(ctypes.c_uint64)()
    n = ctypes.c_uint64(0)
    _capi._check(fn(*args, ctypes.byref(arr), ctypes.byref(n)), _ERRMAP)
    return arr, n.value


# ------------------------------------------------------------------ pi(N)
def prime_count(N: int) -> int:
    """pi(N): the number of primes <= N (2 <= N <= 1e14).

    >>> prime_count(10**9)
    50847534
    """
    _need_int("N", N)
    if N < 2:
        raise VossInvalidNError(f"N must be >= 2, got {N}")
    return _call_u64(_L.voss_primes_prime_count, ctypes.c_uint64(N))


# ------------------------------------------------------------------ result types
class LargeGap:
    """A gap of at least 500 between consecutive primes; position is the larger prime."""

    def __init__(self, position: int, gap: int):
        self.position = position
        self.gap = gap

    def __repr__(self):
        return f"LargeGap(position={self.position:,}, gap={self.gap})"

    def __eq__(self, other):
        if not isinstance(other, LargeGap):
            return NotImplemented
        return (self.position, self.gap) == (other.position, other.gap)


class ChebyshevBias:
    """pi(N;4,1) and pi(N;4,3) with the reference convention (5 in class 1, 3 in class 3)."""

    def __init__(self, pi_4_1, pi_4_3):
        self.pi_4_1 = pi_4_1
        self.pi_4_3 = pi_4_3

    @property
    def difference(self) -> int:
        """pi_4_3 - pi_4_1."""
        return self.pi_4_3 - self.pi_4_1

    def __repr__(self):
        return (f"ChebyshevBias(pi_4_1={self.pi_4_1:,}, pi_4_3={self.pi_4_3:,}, "
                f"difference={self.difference:+,})")


class Statistics:
    """Moments of the consecutive-prime gap distribution (kurtosis is excess)."""

    def __init__(self, mean_gap, std_dev, skewness, kurtosis, total_gaps):
        self.mean_gap = mean_gap
        self.std_dev = std_dev
        self.skewness = skewness
        self.kurtosis = kurtosis
        self.total_gaps = total_gaps

    def __repr__(self):
        return (f"Statistics(mean_gap={self.mean_gap:.4f}, std_dev={self.std_dev:.4f}, "
                f"skewness={self.skewness:.4f}, kurtosis={self.kurtosis:.4f}, "
                f"total_gaps={self.total_gaps})")


# ------------------------------------------------------------------ Context
class Context:
    """Handle for repeated queries on one N.  The first statistics query runs
    the whole sweep on the GPU; later queries read the cache.

    profile: "minimal" (pi only), "standard" or "full" (identical), or the int code.
    """

    def __init__(self, N: int, profile="standard"):
        _need_int("N", N)
        if N < 2:
            raise VossInvalidNError(f"N must be >= 2, got {N}")
        if isinstance(profile, str):
            if profile not in _PROFILES:
                raise ValueError(f"profile must be one of {list(_PROFILES)} or an int; got {profile!r}")
            code = _PROFILES[profile]
        elif isinstance(profile, int):
            code = profile
        else:
            raise TypeError(f"profile must be str or int, got {type(profile).__name__}")
        handle = ctypes.c_void_p()
        _capi._check(_L.voss_primes_ctx_new(ctypes.c_uint64(N), ctypes.c_int(code),
                                            ctypes.byref(handle)), _ERRMAP)
        self._N = N
        self._profile = code
        self._ctx = handle
        self._cache = {}
        self._progress_cb_ref = None

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, tb):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def close(self):
        """Release the handle (idempotent)."""
        if self._ctx is not None:
            _L.voss_primes_ctx_free(self._ctx)
            self._ctx = None

    def _live(self):
        if self._ctx is None:
            raise VossError("Context is closed")
        return self._ctx

    @property
    def N(self) -> int:
        return self._N

    @property
    def profile(self) -> str:
        for name, code in _PROFILES.items():
            if code == self._profile:
                return name
        return str(self._profile)

    def set_progress(self, callback):
        """callback(seg, total, primes_in_seg, cumulative) after each internal
        segment; exceptions inside it are ignored.  None clears it."""
        ctx = self._live()
        if callback is None:
            _L.voss_primes_ctx_set_progress(ctx, ctypes.cast(None, _capi.PROGRESS_CB), None)
            self._progress_cb_ref = None
            return
        if not callable(callback):
            raise TypeError("callback must be callable or None")

        def tramp(seg, total, in_seg, cumulative, _user):
            try:
                callback(seg, total, in_seg, cumulative)
            except Exception:
                pass

        cb = _capi.PROGRESS_CB(tramp)
        _capi._check(_L.voss_primes_ctx_set_progress(ctx, cb, None), {})
        self._progress_cb_ref = cb

    def _cached(self, key, compute):
        ctx = self._live()
        if key not in self._cache:
            self._cache[key] = compute(ctx)
        return self._cache[key]

    def prime_count(self) -> int:
        """pi(N)."""
        return self._cached("pi", lambda c: _call_u64(_L.voss_primes_ctx_prime_count, c))

    def twins(self) -> int:
        """Consecutive primes with gap 2 up to N (STANDARD/FULL)."""
        return self._cached("twins", lambda c: _call_u64(_L.voss_primes_ctx_twins, c))

    def cousin(self) -> int:
        """Consecutive primes with gap 4 up to N (STANDARD/FULL)."""
        return self._cached("cousin", lambda c: _call_u64(_L.voss_primes_ctx_cousin, c))

    def sexy(self) -> int:
        """Consecutive primes with gap 6 up to N (STANDARD/FULL)."""
        return self._cached("sexy", lambda c: _call_u64(_L.voss_primes_ctx_sexy, c))

    def sophie_germain(self) -> int:
        """Primes p <= N with 2p+1 prime (STANDARD/FULL)."""
        return self._cached("sophie", lambda c: _call_u64(_L.voss_primes_ctx_sophie_germain, c))

    def chebyshev(self) -> ChebyshevBias:
        """pi(N;4,1), pi(N;4,3) (STANDARD/FULL)."""
        def get(c):
            a, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
            _capi._check(_L.voss_primes_ctx_chebyshev_bias(c, ctypes.byref(a), ctypes.byref(b)), _ERRMAP)
            return ChebyshevBias(pi_4_1=a.value, pi_4_3=b.value)
        return self._cached("cheb", get)

    def statistics(self) -> Statistics:
        """Mean, standard deviation, skewness, excess kurtosis and count of gaps."""
        def get(c):
            s = _capi.VossPrimesStats()
            _capi._check(_L.voss_primes_ctx_statistics(c, ctypes.byref(s)), _ERRMAP)
            return Statistics(s.mean_gap, s.std_dev, s.skewness, s.kurtosis, s.total_gaps)
        return self._cached("stats", get)

    def large_gaps(self) -> list:
        """Gaps >= 500 as LargeGap, largest first."""
        def get(c):
            n = _call_u64(_L.voss_primes_ctx_large_gaps_count, c)
            out = []
            for i in range(n):
                pos, gap = ctypes.c_uint64(0), ctypes.c_uint32(0)
                _capi._check(_L.voss_primes_ctx_large_gaps_get(c, ctypes.c_uint64(i), ctypes.byref(pos),
                                                               ctypes.byref(gap)), _ERRMAP)
                out.append(LargeGap(position=pos.value, gap=gap.value))
            return out
        return self._cached("large", get)

    def export_csv(self, prefix: str) -> list:
        """Write <prefix>_chebyshev.csv, _large_gaps.csv and _stats.csv; return their paths."""
        ctx = self._live()
        if not isinstance(prefix, str):
            raise TypeError("prefix must be str")
        _capi._check(_L.voss_primes_ctx_export_csv(ctx, prefix.encode("utf-8")), _ERRMAP)
        return [prefix + "_chebyshev.csv", prefix + "_large_gaps.csv", prefix + "_stats.csv"]


# ------------------------------------------------------------------ module functions
def in_range(a: int, b: int) -> list:
    """Primes in [a, b] ascending, at most 1e7 of them (for b > 100 the wheel
    path omits 2, 3 and 5, as the reference does)."""
    if not isinstance(a, int) or not isinstance(b, int):
        raise TypeError("a and b must be int")
    if b < a:
        return []
    arr, n = _call_array(_L.voss_primes_in_range, ctypes.c_uint64(a), ctypes.c_uint64(b))
    return _capi.take_u64(arr, n)


def _in_range_np(a: int, b: int):
    """in_range as a numpy uint64 array (same primes, no per-element list)."""
    import numpy as np
    if b < a:
        return np.zeros(0, dtype=np.uint64)
    arr, n = _call_array(_L.voss_primes_in_range, ctypes.c_uint64(a), ctypes.c_uint64(b))
    if n == 0 or not arr:
        return np.zeros(0, dtype=np.uint64)
    try:
        return np.ctypeslib.as_array(arr, shape=(n,)).copy()
    finally:
        _L.voss_free(arr)


def in_range_with_limit(a: int, b: int, max_count: int) -> list:
    """in_range with a caller-chosen maximum count."""
    if not isinstance(a, int) or not isinstance(b, int):
        raise TypeError("a and b must be int")
    if not isinstance(max_count, int) or max_count <= 0:
        raise ValueError("max_count must be positive int")
    if b < a:
        return []
    arr, n = _call_array(_L.voss_primes_in_range_with_limit, ctypes.c_uint64(a), ctypes.c_uint64(b),
                         ctypes.c_uint64(max_count))
    return _capi.take_u64(arr, n)


def nth(n: int) -> int:
    """The n-th prime, 1-indexed (n <= 5e7)."""
    if not isinstance(n, int):
        raise TypeError("n must be int")
    if n < 1:
        raise VossInvalidNError(f"n must be >= 1 (1-indexed), got {n}")
    return _call_u64(_L.voss_primes_nth, ctypes.c_uint64(n))


def next_prime(x: int) -> int:
    """Smallest prime > x."""
    if not isinstance(x, int):
        raise TypeError("x must be int")
    return _call_u64(_L.voss_primes_next, ctypes.c_uint64(x))


def prev_prime(x: int) -> int:
    """Largest prime < x (x > 2)."""
    if not isinstance(x, int):
        raise TypeError("x must be int")
    return _call_u64(_L.voss_primes_prev, ctypes.c_uint64(x))


def is_prime(x: int) -> bool:
    """Deterministic primality of x (negative x is not prime)."""
    _need_int("x", x)
    if x < 0:
        return False
    return _call_flag(_L.voss_primes_is_prime, ctypes.c_uint64(x))


def factorize(x: int) -> list:
    """Prime factors of x >= 2 with multiplicity, ascending."""
    _need_int("x", x)
    if x < 2:
        raise VossInvalidNError(f"x must be >= 2, got {x}")
    arr, n = _call_array(_L.voss_primes_factorize, ctypes.c_uint64(x))
    return _capi.take_u64(arr, n)


def is_mersenne_prime(p: int) -> bool:
    """Is 2^p - 1 prime?  (2 <= p <= 63)"""
    _need_int("p", p)
    if p < 2:
        raise VossInvalidNError(f"p must be >= 2, got {p}")
    if p > 63:
        raise VossError(f"p must be <= 63, got {p}")
    return _call_flag(_L.voss_primes_is_mersenne_prime, ctypes.c_uint32(p))


def is_fermat_prime(n: int) -> bool:
    """Is 2^(2^n) + 1 prime?  (0 <= n <= 5)"""
    _need_int("n", n)
    if n < 0:
        raise VossInvalidNError(f"n must be >= 0, got {n}")
    if n > 5:
        raise VossError(f"n must be <= 5, got {n}")
    return _call_flag(_L.voss_primes_is_fermat_prime, ctypes.c_uint32(n))


def goldbach_count(n: int) -> int:
    """Number of pairs p1 <= p2 of primes with p1 + p2 = n (even n >= 4, n <= 1e8)."""
    _need_int("n", n)
    if n < 4 or n % 2:
        raise VossInvalidNError(f"n must be even and >= 4, got {n}")
    return _call_u64(_L.voss_primes_goldbach_count, ctypes.c_uint64(n))


def goldbach_partitions(n: int) -> list:
    """The pairs (p1, p2), p1 <= p2, ascending in p1."""
    _need_int("n", n)
    if n < 4 or n % 2:
        raise VossInvalidNError(f"n must be even and >= 4, got {n}")
    arr, pairs = _call_array(_L.voss_primes_goldbach_partitions, ctypes.c_uint64(n))
    flat = _capi.take_u64(arr, 2 * pairs)
    return [(flat[2 * i], flat[2 * i + 1]) for i in range(pairs)]
