"""Empirical tests of prime conjectures up to N (same functions and result
dictionaries as bindings/python/voss/conjectures.py).  They never prove
anything: "no counterexample found" is evidence up to N only.

test_goldbach_upto / iter_goldbach_upto run on the GPU
(voss_primes_goldbach_test_upto); the others are loops over next_prime and
in_range, as in the reference.
"""

import ctypes
import math

from . import _capi, primes

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


def test_bertrand_upto(N):
    """For every n in [2, N] (N <= 1e7): next_prime(n) < 2n."""
    _validate_N(N, 2, 10**7)
    bad = []
    for n in range(2, N + 1):
        if primes.next_prime(n) >= 2 * n:
            bad.append(n)
            if len(bad) >= 10:
                break
    return _result("bertrand", N, bad)


def test_legendre_upto(N):
    """For every n with (n+1)^2 <= N (N <= 1e14): a prime lies in (n^2, (n+1)^2)."""
    _validate_N(N, 2, 10**14)
    bad = []
    for n in range(1, math.isqrt(N)):
        upper = (n + 1) * (n + 1)
        if upper > N:
            break
        if primes.next_prime(n * n) >= upper:
            bad.append(n)
            if len(bad) >= 10:
                break
    return _result("legendre", N, bad)


def test_goldbach_upto(N):
    """Every even n in [4, N] (N <= 1e9) is a sum of two primes (GPU)."""
    _validate_N(N, 4, 10**9)
    return _result("goldbach", N, _goldbach_counterexamples(N))


def test_chebyshev_bias_upto(N):
    """Strict Chebyshev bias pi(x;4,3) > pi(x;4,1) for every prime x <= N (N <= 1e7)."""
    _validate_N(N, 3, 10**7)
    bad = []
    c1 = c3 = 0
    for p in primes.in_range(2, N):
        if p == 2:
            continue
        if p % 4 == 1:
            c1 += 1
        else:
            c3 += 1
        if c3 <= c1:
            bad.append(p)
            if len(bad) >= 10:
                break
    return _result("chebyshev_bias", N, bad)


def iter_bertrand_upto(N):
    """Yield (n, next_prime(n)) for n in [2, N]."""
    _validate_N(N, 2, 10**7)
    for n in range(2, N + 1):
        yield n, primes.next_prime(n)


def iter_legendre_upto(N):
    """Yield (n, next_prime(n^2)) while (n+1)^2 <= N."""
    _validate_N(N, 2, 10**14)
    for n in range(1, math.isqrt(N)):
        if (n + 1) * (n + 1) > N:
            break
        yield n, primes.next_prime(n * n)


def iter_chebyshev_bias_upto(N):
    """Yield (p, c1, c3) for each odd prime p <= N."""
    _validate_N(N, 3, 10**7)
    c1 = c3 = 0
    for p in primes.in_range(2, N):
        if p == 2:
            continue
        if p % 4 == 1:
            c1 += 1
        else:
            c3 += 1
        yield p, c1, c3


def iter_goldbach_upto(N):
    """Yield the even n in [4, N] with no partition (GPU; normally none)."""
    _validate_N(N, 4, 10**9)
    yield from _goldbach_counterexamples(N)


__all__ = ["test_bertrand_upto", "test_legendre_upto", "test_goldbach_upto",
           "test_chebyshev_bias_upto", "iter_bertrand_upto", "iter_legendre_upto",
           "iter_goldbach_upto", "iter_chebyshev_bias_upto"]
