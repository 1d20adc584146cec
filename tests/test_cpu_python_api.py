"""The Python package's contract that needs no device (the reference's
bindings/python/voss/primes.py behaviour): argument checks before any call
into the library, result types and their reprs, contexts below 7 (answered on
the host), closed contexts, the host-side number theory and error mapping."""

import pytest


def test_version_and_modules(voss):
    assert voss.__version__ == "0.7.0"
    for name in ("primes", "conjectures", "numbers", "stats", "serve"):
        assert hasattr(voss, name)


def test_argument_checks_before_the_library(voss):
    P = voss.primes
    with pytest.raises(TypeError):
        P.prime_count("10")
    with pytest.raises(TypeError):
        P.prime_count(10.0)
    with pytest.raises(P.VossInvalidNError):
        P.prime_count(1)
    with pytest.raises(P.VossInvalidNError):
        P.prime_count(True)  # bool is an int (1)
    with pytest.raises(P.VossInvalidNError):
        P.Context(1)
    with pytest.raises(ValueError):
        P.Context(100, profile="bogus")
    with pytest.raises(TypeError):
        P.Context(100, profile=1.5)
    assert P.in_range(10, 5) == []
    assert P.is_prime(-7) is False
    with pytest.raises(TypeError):
        P.nth("3")
    with pytest.raises(P.VossInvalidNError):
        P.nth(0)


def test_small_contexts_on_the_host(voss):
    # context.cu:228-247: N < 7 is answered without the sieve
    P = voss.primes
    for N, pi, twins, cousin in [(2, 1, 0, 0), (3, 2, 0, 0), (4, 2, 0, 0), (5, 3, 1, 0), (6, 3, 1, 0)]:
        with P.Context(N) as c:
            assert c.prime_count() == pi
            assert c.twins() == twins
            assert c.cousin() == cousin
            assert c.large_gaps() == []
            b = c.chebyshev()
            assert (b.pi_4_1, b.pi_4_3) == (0, 0)
            assert c.prime_count() is c.prime_count()  # cached object
    with P.Context(5, profile="minimal") as c:
        assert c.profile == "minimal"
        assert c.prime_count() == 3
        with pytest.raises(P.VossError):
            c.twins()


def test_closed_context(voss):
    P = voss.primes
    c = P.Context(5)
    c.close()
    c.close()  # idempotent
    with pytest.raises(P.VossError, match="Context is closed"):
        c.prime_count()


def test_result_reprs(voss):
    P = voss.primes
    assert repr(P.LargeGap(738832928467, 540)) == "LargeGap(position=738,832,928,467, gap=540)"
    assert P.LargeGap(1, 2) == P.LargeGap(1, 2) and P.LargeGap(1, 2) != P.LargeGap(1, 4)
    b = P.ChebyshevBias(25423491, 25424042)
    assert b.difference == 551
    assert repr(b) == "ChebyshevBias(pi_4_1=25,423,491, pi_4_3=25,424,042, difference=+551)"
    s = P.Statistics(19.6666, 1.5, 2.25, 3.125, 50847533)
    assert repr(s).startswith("Statistics(mean_gap=19.6666, std_dev=1.5000, skewness=2.2500, kurtosis=3.1250")


def test_host_number_theory(voss):
    N = voss.numbers
    assert N.phi(36) == 12 and N.tau(36) == 9 and N.sigma(36) == 91 and N.mu(30) == -1 and N.mu(12) == 0
    assert N.gcd(84, 36) == 12 and N.lcm(4, 6) == 12
    assert N.fibonacci(93) == 12200160415121876738 and N.factorial(20) == 2432902008176640000
    with pytest.raises(voss.VossError):
        N.fibonacci(94)
    P = voss.primes
    assert P.factorize(2**2 * 3 * 1000003) == [2, 2, 3, 1000003]
    assert P.is_mersenne_prime(61) is True and P.is_mersenne_prime(59) is False
    assert P.is_fermat_prime(4) is True and P.is_fermat_prime(5) is False
    # host Miller-Rabin and next/prev need no device (reference core/miller_rabin.cpp)
    assert [P.is_prime(x) for x in (0, 1, 2, 3, 4, 5, 100, 7919, 1000000007)] == \
        [False, False, True, True, False, True, False, True, True]
    assert P.next_prime(10**9) == 1000000007 and P.prev_prime(10**9) == 999999937
