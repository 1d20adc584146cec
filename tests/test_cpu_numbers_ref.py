"""The host-side number functions (voss_numbers_*, factorize, Mersenne and
Fermat tests) against the reference's own C++ for them, compiled from its
sources into oracle/_ref (library/src/numbers/*.cpp, primes/factorize.cpp,
primes/mersenne_fermat.cpp): return code and value for every n up to 1e5 and
for seeded random 64-bit inputs, plus the reference test file's known values
(library/tests/test_numbers.py) through the Python package."""

import ctypes
import random

import pytest

U64, P64 = ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)


def _ref():
    from oracle import oracle as O
    r = O.ref_cpu()
    if r is None:
        pytest.skip("reference CPU library not built (reference tree absent)")
    return r


def _bind(L):
    for f in ("phi", "tau", "sigma", "fibonacci", "factorial"):
        fn = getattr(L, "voss_numbers_" + f)
        fn.argtypes, fn.restype = [U64, P64], ctypes.c_int
    L.voss_numbers_mu.argtypes, L.voss_numbers_mu.restype = [U64, ctypes.POINTER(ctypes.c_int)], ctypes.c_int
    for f in ("gcd", "lcm"):
        fn = getattr(L, "voss_numbers_" + f)
        fn.argtypes, fn.restype = [U64, U64, P64], ctypes.c_int
    L.voss_primes_factorize.argtypes = [U64, ctypes.POINTER(P64), P64]
    L.voss_primes_factorize.restype = ctypes.c_int
    L.voss_primes_is_mersenne_prime.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]
    L.voss_primes_is_fermat_prime.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]
    L.voss_free.argtypes = [ctypes.c_void_p]
    return L


@pytest.fixture(scope="module")
def libs(voss):
    from voss_amd import _capi
    return _bind(_capi._lib), _bind(_ref())


def _u(L, f, *args):
    out = U64(0xA5A5A5A5)
    rc = getattr(L, f)(*args, ctypes.byref(out))
    return rc, (out.value if rc == 0 else None)


def _mu(L, n):
    out = ctypes.c_int(7)
    rc = L.voss_numbers_mu(n, ctypes.byref(out))
    return rc, (out.value if rc == 0 else None)


def _factors(L, n):
    arr, cnt = P64(), U64()
    rc = L.voss_primes_factorize(n, ctypes.byref(arr), ctypes.byref(cnt))
    if rc != 0:
        return rc, None
    v = [arr[i] for i in range(cnt.value)]
    if arr:
        L.voss_free(arr)
    return rc, v


def _inputs():
    # The reference's Pollard rho adds residues as (a + b) % m in 64 bits
    # (factorize.cpp:31-33), which wraps for m >= 2^63, so on a composite
    # cofactor above 2^63 it may never stop: the comparison stays below 2^63
    # except for inputs that its trial division or Miller-Rabin settles.
    rng = random.Random(20261204)
    ns = list(range(0, 100001))
    ns += [rng.getrandbits(63) for _ in range(300)]
    ns += [rng.getrandbits(rng.randint(33, 62)) for _ in range(300)]
    # semiprimes of two ~31-bit primes, prime powers and the top of the range
    ns += [2147483647 * 2147483629, 2**61 - 1, (2**31 - 1) ** 2, 3**40, 2**63, 2**64 - 1, 2**64 - 59,
           999999999989 * 1000003]
    return ns


def test_multiplicative_functions_vs_reference(libs):
    ours, ref = libs
    for n in _inputs():
        for f in ("voss_numbers_phi", "voss_numbers_tau", "voss_numbers_sigma"):
            assert _u(ours, f, n) == _u(ref, f, n), (f, n)
        assert _mu(ours, n) == _mu(ref, n), ("mu", n)


def test_factorize_vs_reference(libs):
    ours, ref = libs
    for n in _inputs():
        assert _factors(ours, n) == _factors(ref, n), n


def test_factorize_above_2_63(libs, oracle):
    # where the reference may not terminate: the factors must be primes whose
    # product is n, in ascending order
    ours, _ = libs
    rng = random.Random(11)
    ns = [4294967291 * 4294967279, 4294967279 * 4294967279, 18446744073709551557 - 2]
    ns += [rng.getrandbits(64) | (1 << 63) for _ in range(200)]
    for n in ns:
        rc, f = _factors(ours, n)
        assert rc == 0 and f == sorted(f), n
        prod = 1
        for p in f:
            assert oracle.is_prime(p), (n, p)
            prod *= p
        assert prod == n


def test_gcd_lcm_vs_reference(libs):
    ours, ref = libs
    rng = random.Random(7)
    pairs = [(a, b) for a in range(0, 60) for b in range(0, 60)]
    pairs += [(rng.getrandbits(64), rng.getrandbits(64)) for _ in range(2000)]
    pairs += [(rng.getrandbits(32), rng.getrandbits(33)) for _ in range(2000)]
    pairs += [(2**64 - 1, 2**63), (2**32, 2**32 + 1), (2**63, 2), (0, 2**64 - 1)]
    for a, b in pairs:
        for f in ("voss_numbers_gcd", "voss_numbers_lcm"):
            assert _u(ours, f, a, b) == _u(ref, f, a, b), (f, a, b)


def test_fibonacci_factorial_mersenne_fermat_vs_reference(libs):
    ours, ref = libs
    for n in list(range(0, 120)) + [2**32, 2**64 - 1]:
        for f in ("voss_numbers_fibonacci", "voss_numbers_factorial"):
            assert _u(ours, f, n) == _u(ref, f, n), (f, n)
    for p in range(0, 80):
        a, b = ctypes.c_int(9), ctypes.c_int(9)
        ra = ours.voss_primes_is_mersenne_prime(p, ctypes.byref(a))
        rb = ref.voss_primes_is_mersenne_prime(p, ctypes.byref(b))
        assert (ra, a.value if ra == 0 else None) == (rb, b.value if rb == 0 else None), p
    for n in range(0, 12):
        a, b = ctypes.c_int(9), ctypes.c_int(9)
        ra = ours.voss_primes_is_fermat_prime(n, ctypes.byref(a))
        rb = ref.voss_primes_is_fermat_prime(n, ctypes.byref(b))
        assert (ra, a.value if ra == 0 else None) == (rb, b.value if rb == 0 else None), n


def test_reference_number_values(voss):
    # library/tests/test_numbers.py, the explicit values
    N = voss.numbers
    assert [N.tau(n) for n in (1, 2, 4, 6, 12, 100, 360, 97, 1000000007, 2**10, 3**5)] == \
        [1, 2, 3, 4, 6, 9, 24, 2, 2, 11, 6]
    assert [N.sigma(n) for n in (1, 2, 6, 12, 28, 100, 360, 97, 1000000007, 2**10, 3**5)] == \
        [1, 3, 12, 28, 56, 217, 1170, 98, 1000000008, 2047, 364]
    assert [N.mu(n) for n in (1, 2, 3, 6, 30, 210, 12, 4, 9, 100, 97, 1000000007)] == \
        [1, -1, -1, 1, -1, 1, 0, 0, 0, 0, -1, -1]
    assert [N.gcd(a, b) for a, b in ((12, 18), (48, 18), (17, 5), (100, 75), (0, 0), (5, 0), (0, 7))] == \
        [6, 6, 1, 25, 0, 5, 7]
    assert [N.lcm(a, b) for a, b in ((4, 6), (21, 6), (7, 11), (0, 5), (7, 0), (0, 0))] == [12, 42, 77, 0, 0, 0]
    assert N.fibonacci(92) == 7540113804746346429 and N.fibonacci(93) == 12200160415121876738
    assert [N.factorial(n) for n in (0, 1, 5, 10, 20)] == [1, 1, 120, 3628800, 2432902008176640000]
    for bad in (lambda: N.phi(0), lambda: N.tau(0), lambda: N.sigma(0), lambda: N.fibonacci(94),
                lambda: N.factorial(21), lambda: N.lcm(2**63, 3)):
        with pytest.raises(voss.VossError):
            bad()
    for bad in (lambda: N.phi(3.5), lambda: N.phi("12"), lambda: N.gcd(1.5, 2)):
        with pytest.raises(TypeError):
            bad()
