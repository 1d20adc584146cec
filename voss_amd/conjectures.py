"""Empirical tests of prime conjectures up to N (same functions and result
dictionaries as bindings/python/voss/conjectures.py).  They never prove
anything: "no counterexample found" is evidence up to N only.

Where the reference loops over next_prime / in_range one value at a time
(library/bindings/python/voss/conjectures.py), these run on the GPU in bulk:
Bertrand and the Chebyshev bias from one prime list, Legendre from one
batched next-prime launch over n^2, Goldbach through
voss_primes_goldbach_test_upto.  The answers (and the first ten
counterexamples, ascending) are the ones the per-value loops give.
"""

import ctypes
import math

import numpy as np

from . import _capi, primes, stats

_NOTE = "Empirical test up to N. Does NOT prove the conjecture."


def _result(name, N, found):
    return {
        "conjecture": name,
        "N": N,
        "status": "counterexample_found" if found else "no_counterexample_found_upto_N",
        "counterexamples": list(found[:10]),
        "note": _NOTE,
    }


def _validate_N(N, min_N, max_N):
    if not isinstance(N, int):
        raise TypeError(f"N must be int, got {type(N).__name__}")
    if N < min_N:
        raise primes.VossInvalidNError(f"N must be >= {min_N} for this test, got {N}")
    if N > max_N:
        raise primes.VossInvalidArgError(f"N must be <= {max_N} for this test, got {N}.")


def _goldbach_counterexamples(N, limit=10):
    arr = ctypes.POINTER(ctypes.c_uint64)()
    cnt = ctypes.c_uint64(0)
    rc = _capi._lib.voss_primes_goldbach_test_upto(ctypes.c_uint64(N), ctypes.byref(arr),
                                                   ctypes.byref(cnt), ctypes.c_uint64(limit))
    _capi._check(rc, {})
    return _capi.take_u64(arr, cnt.value)


def _prime_array(hi):
    """All primes <= hi as a uint64 array (the GPU list; 2, 3, 5 prepended)."""
    head = np.array([q for q in (2, 3, 5) if q <= hi], dtype=np.uint64)
    if hi < 7:
        return head
    return np.concatenate([head, primes._in_range_np(7, hi)])


def _next_primes(lo, hi, P):
    """next_prime(n) for n in [lo, hi] from an ascending prime array P reaching past hi."""
    n = np.arange(lo, hi + 1, dtype=np.uint64)
    return n, P[np.searchsorted(P, n, side="right")]


def test_bertrand_upto(N):
    """For every n in [2, N] (N <= 1e7): next_prime(n) < 2n."""
    _validate_N(N, 2, 10**7)
    n, nxt = _next_primes(2, N, _prime_array(2 * N + 2))
    bad = n[nxt >= 2 * n][:10]
    return _result("bertrand", N, [int(v) for v in bad])


def _legendre_next(N):
    """(n, next_prime(n^2)) for n = 1 .. isqrt(N) - 1 (batched GPU next-prime)."""
    n = np.arange(1, math.isqrt(N), dtype=np.uint64)
    nxt = np.empty_like(n)
    step = 1 << 20
    for i in range(0, len(n), step):
        sq = n[i:i + step] * n[i:i + step]
        nxt[i:i + step] = stats.next_prime_batch(sq, as_numpy=True)
    return n, nxt


def test_legendre_upto(N):
    """For every n with (n+1)^2 <= N (N <= 1e14): a prime lies in (n^2, (n+1)^2)."""
    _validate_N(N, 2, 10**14)
    n, nxt = _legendre_next(N)
    bad = n[nxt >= (n + 1) * (n + 1)][:10]
    return _result("legendre", N, [int(v) for v in bad])


def test_goldbach_upto(N):
    """Every even n in [4, N] (N <= 1e9) is a sum of two primes (GPU)."""
    _validate_N(N, 4, 10**9)
    return _result("goldbach", N, _goldbach_counterexamples(N))


def _chebyshev_counts(N):
    # the reference walks primes.in_range(2, N), which for N > 100 omits 2, 3
    # and 5 (library/src/primes/range.cu:130-134); the same list is used here
    P = primes._in_range_np(2, N)
    P = P[P != 2]
    c1 = np.cumsum(P % 4 == 1, dtype=np.int64)
    c3 = np.arange(1, len(P) + 1, dtype=np.int64) - c1
    return P, c1, c3


def test_chebyshev_bias_upto(N):
    """Strict Chebyshev bias pi(x;4,3) > pi(x;4,1) for every prime x <= N (N <= 1e7)."""
    _validate_N(N, 3, 10**7)
    P, c1, c3 = _chebyshev_counts(N)
    bad = P[c3 <= c1][:10]
    return _result("chebyshev_bias", N, [int(v) for v in bad])


def iter_bertrand_upto(N):
    """Yield (n, next_prime(n)) for n in [2, N]."""
    _validate_N(N, 2, 10**7)
    n, nxt = _next_primes(2, N, _prime_array(2 * N + 2))
    for a, b in zip(n.tolist(), nxt.tolist()):
        yield a, b


def iter_legendre_upto(N):
    """Yield (n, next_prime(n^2)) while (n+1)^2 <= N."""
    _validate_N(N, 2, 10**14)
    n, nxt = _legendre_next(N)
    for a, b in zip(n.tolist(), nxt.tolist()):
        yield a, b


def iter_chebyshev_bias_upto(N):
    """Yield (p, c1, c3) for each odd prime p <= N."""
    _validate_N(N, 3, 10**7)
    P, c1, c3 = _chebyshev_counts(N)
    for a, b, c in zip(P.tolist(), c1.tolist(), c3.tolist()):
        yield a, b, c


def iter_goldbach_upto(N):
    """Yield the even n in [4, N] with no partition (GPU; normally none)."""
    _validate_N(N, 4, 10**9)
    yield from _goldbach_counterexamples(N)


__all__ = ["test_bertrand_upto", "test_legendre_upto", "test_goldbach_upto",
           "test_chebyshev_bias_upto", "iter_bertrand_upto", "iter_legendre_upto",
           "iter_goldbach_upto", "iter_chebyshev_bias_upto"]
