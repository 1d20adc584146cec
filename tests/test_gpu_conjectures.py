"""voss_amd.conjectures against the reference's per-value loops
(library/bindings/python/voss/conjectures.py:48-129, 181-232, 249-328)
restated over the oracle's next_prime and primes_in_range, plus the values,
shapes and errors that the reference's own test file asserts
(library/tests/test_conjectures.py)."""

import math

import pytest

pytestmark = pytest.mark.gpu


def _ref_in_range(oracle, a, b):
    # the reference's in_range omits 2, 3 and 5 when b > 100 (range.cu:130-134)
    ps = oracle.primes_in_range(max(a, 2), b)
    return [p for p in ps if b <= 100 or p > 5]


def _next(oracle, x):
    rc, v = oracle.next_prime(x)
    assert rc == 0
    return v


def _loop_bertrand(oracle, N):
    return [(n, _next(oracle, n)) for n in range(2, N + 1)]


def _loop_legendre(oracle, N):
    out = []
    for n in range(1, math.isqrt(N)):
        if (n + 1) ** 2 > N:
            break
        out.append((n, _next(oracle, n * n)))
    return out


def _loop_chebyshev(oracle, N):
    rows, c1, c3 = [], 0, 0
    for p in _ref_in_range(oracle, 2, N):
        if p == 2:
            continue
        if p % 4 == 1:
            c1 += 1
        else:
            c3 += 1
        rows.append((p, c1, c3))
    return rows


@pytest.mark.parametrize("N", [2, 3, 10, 20, 1000, 30011, 200003])
def test_bertrand_vs_oracle_loop(voss, oracle, N):
    c = voss.conjectures
    want = _loop_bertrand(oracle, N)
    assert list(c.iter_bertrand_upto(N)) == want
    r = c.test_bertrand_upto(N)
    assert r["counterexamples"] == [n for n, p in want if p >= 2 * n][:10] == []
    assert r["status"] == "no_counterexample_found_upto_N"


@pytest.mark.parametrize("N", [2, 3, 4, 99, 100, 101, 10**6, 10**8 + 7, 10**10])
def test_legendre_vs_oracle_loop(voss, oracle, N):
    c = voss.conjectures
    want = _loop_legendre(oracle, N)
    assert list(c.iter_legendre_upto(N)) == want
    assert c.test_legendre_upto(N)["counterexamples"] == [n for n, q in want if q >= (n + 1) ** 2][:10] == []


@pytest.mark.parametrize("N", [3, 4, 5, 10, 50, 100, 101, 1000, 10**5, 10**7])
def test_chebyshev_bias_vs_oracle_loop(voss, oracle, N):
    c = voss.conjectures
    rows = _loop_chebyshev(oracle, N)
    assert list(c.iter_chebyshev_bias_upto(N)) == rows
    bad = [p for p, c1, c3 in rows if c3 <= c1]
    r = c.test_chebyshev_bias_upto(N)
    assert r["counterexamples"] == bad[:10]
    assert r["status"] == ("counterexample_found" if bad else "no_counterexample_found_upto_N")


def test_goldbach_vs_oracle(voss, oracle):
    c = voss.conjectures
    for N in (4, 5, 6, 1000, 10**6):
        assert c.test_goldbach_upto(N)["counterexamples"] == oracle.goldbach_test_upto(N)[1] == []
        assert list(c.iter_goldbach_upto(N)) == []


def test_reference_conjecture_values(voss):
    # library/tests/test_conjectures.py:35-139, 229-276
    c = voss.conjectures
    assert set(c.__all__) == {"test_bertrand_upto", "test_legendre_upto", "test_goldbach_upto",
                              "test_chebyshev_bias_upto", "iter_bertrand_upto", "iter_legendre_upto",
                              "iter_goldbach_upto", "iter_chebyshev_bias_upto"}
    r = c.test_bertrand_upto(10)
    assert set(r) == {"conjecture", "N", "status", "counterexamples", "note"}
    assert r["conjecture"] == "bertrand" and r["N"] == 10 and isinstance(r["counterexamples"], list)
    assert "Empirical test" in r["note"] and "Does NOT prove" in r["note"]
    assert c.test_bertrand_upto(1000)["counterexamples"] == []
    assert c.test_bertrand_upto(2)["status"] == "no_counterexample_found_upto_N"
    r = c.test_legendre_upto(10**4)
    assert r["status"] == "no_counterexample_found_upto_N" and r["counterexamples"] == []
    assert c.test_goldbach_upto(1000)["status"] == "no_counterexample_found_upto_N"
    r = c.test_chebyshev_bias_upto(100)
    assert r["status"] == "counterexample_found" and {5, 17, 41} <= set(r["counterexamples"])
    assert len(c.test_chebyshev_bias_upto(1000)["counterexamples"]) <= 10
    items = list(c.iter_bertrand_upto(20))
    assert len(items) == 19 and all(n < p < 2 * n for n, p in items)
    gen = c.iter_bertrand_upto(10**6)
    assert next(gen) == (2, 3)
    gen.close()
    items = list(c.iter_legendre_upto(100))
    assert len(items) == 9 and all(n * n < p < (n + 1) ** 2 for n, p in items)
    bad = [p for p, c1, c3 in c.iter_chebyshev_bias_upto(50) if c3 <= c1]
    assert {5, 17, 41} <= set(bad)
    assert list(c.iter_goldbach_upto(10**6)) == []


def test_reference_conjecture_errors(voss):
    # library/tests/test_conjectures.py:150-221, 279-298
    c = voss.conjectures
    for f, arg in ((c.test_bertrand_upto, "hello"), (c.test_legendre_upto, 3.14),
                   (c.test_goldbach_upto, None), (c.test_chebyshev_bias_upto, [100])):
        with pytest.raises(TypeError):
            f(arg)
    for f, arg in ((c.test_bertrand_upto, 1), (c.test_bertrand_upto, 10**8), (c.test_legendre_upto, 1),
                   (c.test_goldbach_upto, 3), (c.test_goldbach_upto, 10**10), (c.test_chebyshev_bias_upto, 2),
                   (c.test_chebyshev_bias_upto, 10**8)):
        with pytest.raises(voss.VossError):
            f(arg)
    with pytest.raises(TypeError):
        list(c.iter_bertrand_upto("hello"))
    with pytest.raises(TypeError):
        list(c.iter_legendre_upto(3.14))
    with pytest.raises(voss.VossError):
        list(c.iter_bertrand_upto(10**8))
    with pytest.raises(voss.VossError):
        list(c.iter_goldbach_upto(10**10))
