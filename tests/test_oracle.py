"""Pins the CPU oracle before it is trusted as the checker.

Checks, each chosen so a plausible slip (wrong seed, off-by-one at a segment
or chunk edge, wrong residue class, dropped boundary gap) fails one of them:
  * the reference's own known answers (library/tests/test_primes.py,
    sanitizer_test.cpp, test_prime_count.cpp, library/README.md);
  * brute force on small N with an independent Python sieve;
  * closed-form invariants (sum of the histogram, telescoping sum of gaps,
    class totals);
  * the reference's CPU sources compiled from /root/reference into
    oracle/_ref (Miller-Rabin, nth, next/prev, Goldbach counts).
"""

import ctypes
import random

import pytest


def simple_primes(n):
    if n < 2:
        return []
    s = bytearray([1]) * (n + 1)
    s[0] = s[1] = 0
    for i in range(2, int(n ** 0.5) + 1):
        if s[i]:
            s[i * i::i] = bytearray(len(s[i * i::i]))
    return [i for i in range(n + 1) if s[i]]


def brute_stats(N, cp=0):
    ps = simple_primes(N)
    hist = [0] * 100000
    recs, rmax = [], 0
    for a, b in zip(ps, ps[1:]):
        g = b - a
        hist[g] += 1
        if g > rmax:
            recs.append((b, g))
            rmax = g
    odd = [p for p in ps if p != 2]
    c41 = sum(1 for p in odd if p % 4 == 1)
    c31 = sum(1 for p in ps if p % 3 == 1)
    c32 = sum(1 for p in ps if p % 3 == 2)
    cps = []
    if cp:
        acc = [0, 0, 0, 0, 0]
        it = iter(ps)
        p = next(it, None)
        for j in range(1, N // cp + 1):
            while p is not None and p <= j * cp:
                acc[0] += 1
                if p % 4 == 1:
                    acc[1] += 1
                elif p % 4 == 3:
                    acc[2] += 1
                if p % 3 == 1:
                    acc[3] += 1
                elif p % 3 == 2:
                    acc[4] += 1
                p = next(it, None)
            cps.append(tuple(acc))
    return {"prime_count": len(ps), "hist": hist, "records": recs, "pi_4_1": c41,
            "pi_4_3": len(odd) - c41, "pi_3_1": c31, "pi_3_2": c32, "checkpoints": cps,
            "last_prime": ps[-1]}


# ---------------------------------------------------------------- known answers
@pytest.mark.parametrize("N,pi", [(2, 1), (3, 2), (5, 3), (10, 4), (100, 25), (10**6, 78498),
                                  (10**7, 664579), (10**8, 5761455), (10**9, 50847534)])
def test_known_pi(oracle, N, pi):
    assert oracle.prime_count(N) == pi
    assert oracle.full_stats(N)["prime_count"] == pi


def test_known_gap_counts(oracle):
    s = oracle.full_stats(10**6)
    assert (s["hist"][2], s["hist"][4], s["hist"][6]) == (8169, 8143, 13549)
    assert sum(s["hist"]) == 78497 and s["large_gaps"] == []
    assert (s["pi_4_1"], s["pi_4_3"]) == (39175, 39322)
    s = oracle.full_stats(10**8)
    assert s["hist"][2] == 440312
    assert (s["pi_4_1"], s["pi_4_3"]) == (2880504, 2880950)
    s = oracle.full_stats(10**9)
    assert (s["hist"][2], s["hist"][4], s["hist"][6]) == (3424506, 3424679, 6089791)
    assert s["pi_4_3"] - s["pi_4_1"] == 551
    assert s["large_gaps"] == []
    mean = sum(g * h for g, h in enumerate(s["hist"])) / sum(s["hist"])
    assert round(mean, 4) == 19.6666
    # gap repulsion P(6)/P(2) at 1e9 as the frozen v7 program records it
    # (src/voss_w30_v7.cu:524-534, 4 decimals)
    assert abs(s["hist"][6] / s["hist"][2] - 1.7783) < 5e-5


def test_known_chebyshev_1e4(oracle):
    s = oracle.full_stats(10**4)
    assert (s["pi_4_1"], s["pi_4_3"]) == (609, 619)


def test_small_n_reference_quirks(oracle):
    # context.cu:228-247: below 7 the mod-4 tallies stay 0
    for N, pi in [(2, 1), (3, 2), (4, 2), (5, 3), (6, 3)]:
        s = oracle.full_stats(N)
        assert s["prime_count"] == pi and s["pi_4_1"] == 0 and s["pi_4_3"] == 0
    s = oracle.full_stats(7)
    assert (s["pi_4_1"], s["pi_4_3"]) == (1, 2)  # {5}, {3, 7}


# ---------------------------------------------------------------- brute force
@pytest.mark.parametrize("N", [7, 8, 29, 30, 31, 97, 1000, 65536 * 30 + 11, 3 * 10**6 + 1])
def test_brute_force(oracle, N):
    cp = 1000 if N >= 10**5 else 0
    o = oracle.full_stats(N, cp_interval=cp)
    b = brute_stats(N, cp)
    o_hist = list(o["hist"])
    assert o["prime_count"] == b["prime_count"]
    assert o_hist == b["hist"]
    assert o["records"] == b["records"]
    assert (o["pi_4_1"], o["pi_4_3"]) == (b["pi_4_1"], b["pi_4_3"])
    assert (o["pi_3_1"], o["pi_3_2"]) == (b["pi_3_1"], b["pi_3_2"])
    assert o["checkpoints"] == b["checkpoints"]
    assert o["last_prime"] == b["last_prime"]


def test_range_stats_brute(oracle):
    rng = random.Random(1)
    ps = simple_primes(2 * 10**6)
    for _ in range(60):
        a = rng.randint(0, 10**6)
        b = a + rng.randint(0, 10**6)
        if rng.random() < 0.2:
            a = rng.randint(0, 12)
        q = [p for p in ps if a <= p <= b]
        h = [0] * 257
        for x, y in zip(q, q[1:]):
            g = y - x
            h[g // 2 if (g % 2 == 0 and 2 <= g <= 512) else 0] += 1
        r = oracle.range_stats(a, b)
        assert (r["count"], r["first"], r["last"]) == (len(q), q[0] if q else 0, q[-1] if q else 0)
        assert r["hist"] == h
        assert oracle.primes_in_range(a, b) == q


# ---------------------------------------------------------------- invariants at scale
@pytest.mark.parametrize("N", [10**9, 2 * 10**9 + 11])
def test_invariants(oracle, N):
    s = oracle.full_stats(N, cp_interval=10**8)
    h = s["hist"]
    assert sum(h) == s["prime_count"] - 1
    assert sum(g * c for g, c in enumerate(h)) == s["last_prime"] - 2  # telescoping
    assert s["pi_4_1"] + s["pi_4_3"] + 1 == s["prime_count"]
    assert s["pi_3_1"] + s["pi_3_2"] + 1 == s["prime_count"]
    assert s["max_gap"] == max(g for g, c in enumerate(h) if c)
    gaps = [g for _, g in s["records"]]
    assert gaps == sorted(set(gaps)) and gaps[-1] == s["max_gap"]
    # checkpoints agree with independent counts at their x
    for j, cp in enumerate(s["checkpoints"][:3], start=1):
        assert cp[0] == oracle.prime_count(j * 10**8)


def test_records_known(oracle):
    # maximal gaps (A002386 / A005250): gap 282 after 436273009, the largest below 1e9
    s = oracle.full_stats(10**9)
    assert s["records"][-1] == (436273009 + 282, 282)
    assert [g for _, g in s["records"][:12]] == [1, 2, 4, 6, 8, 14, 18, 20, 22, 34, 36, 44]


# ---------------------------------------------------------------- against the reference's CPU code
def _ref():
    from oracle import oracle as O
    r = O.ref_cpu()
    if r is None:
        pytest.skip("reference CPU library not built (reference tree absent)")
    return r


def test_mr_vs_reference_cpu(oracle):
    r = _ref()
    r.voss_primes_is_prime.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]
    rng = random.Random(2)
    xs = list(range(0, 2000)) + [2047, 3215031751, 3825123056546413051, 2**64 - 59, 2**64 - 1]
    xs += [rng.getrandbits(64) for _ in range(2000)]
    for x in xs:
        out = ctypes.c_int()
        r.voss_primes_is_prime(x, ctypes.byref(out))
        assert bool(out.value) == oracle.is_prime(x), x


def test_nth_next_prev_vs_reference_cpu(oracle):
    r = _ref()
    for f in ("voss_primes_nth", "voss_primes_next", "voss_primes_prev"):
        getattr(r, f).argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    out = ctypes.c_uint64()
    for n in [1, 2, 3, 10, 100, 1000, 99999, 100000, 100001, 10**6, 3 * 10**6]:
        assert r.voss_primes_nth(n, ctypes.byref(out)) == 0
        assert out.value == oracle.nth_prime(n)
    rng = random.Random(4)
    for x in [0, 1, 2, 3, 4, 10, 1000, 10**9] + [rng.getrandbits(50) for _ in range(200)]:
        assert r.voss_primes_next(x, ctypes.byref(out)) == 0
        assert out.value == oracle.next_prime(x)[1]
        if x > 2:
            assert r.voss_primes_prev(x, ctypes.byref(out)) == 0
            assert out.value == oracle.prev_prime(x)[1]


def test_goldbach_vs_reference_cpu(oracle):
    r = _ref()
    r.voss_primes_goldbach_count.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    out = ctypes.c_uint64()
    ns = [4, 6, 8, 10, 100, 1000, 10**4, 10**5, 2 * 10**5 + 2]
    for n in ns:
        r.voss_primes_goldbach_count(n, ctypes.byref(out))
        assert out.value == oracle.goldbach_count_mr(n) == oracle.goldbach_count_sieve(n)
    assert oracle.goldbach_counts(ns) == [oracle.goldbach_count_sieve(n) for n in ns]
    assert [oracle.goldbach_count_sieve(n) for n in (10, 100, 1000, 10**4, 10**5)] == [2, 6, 28, 127, 810]


def test_goldbach_upto_and_sophie(oracle):
    assert oracle.goldbach_test_upto(10**6) == (0, [])
    # planted defect: partners restricted to p <= pmax give counterexamples a
    # brute force lists exactly (count, order, the first max_ce, block edges)
    ps = set(simple_primes(3 * 10**6))
    for N, pmax, cap in [(2000, 13, 10), (3 * 10**6, 31, 40), (3 * 10**6, 7, 10**6)]:
        want = [n for n in range(6, N + 1, 2)
                if not any(p in ps and (n - p) in ps for p in range(3, min(n // 2, pmax) + 1, 2))]
        got_n, got = oracle.goldbach_test_upto(N, max_ce=cap, pmax=pmax)
        assert got_n == len(want) and got == want[:cap], (N, pmax)
    assert [oracle.sophie_germain(n) for n in (10, 30, 1000, 100, 10**6)] == [3, 6, 37, 10, 7746]


def test_reference_large_gap_rule(oracle):
    # the reference's large list drops gaps that straddle its 5e10+10 segments;
    # the rule keeps gaps inside one segment and drops straddling ones
    assert oracle.reference_large_gaps([(50000000030, 40)], 10**11) == []
    assert oracle.reference_large_gaps([(50000000009, 10)], 10**11) == [(50000000009, 10)]
