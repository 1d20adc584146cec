"""GPU parity: the HIP path (through the C ABI / Python package) against the
CPU oracle on identical inputs, bit for bit.

The oracle itself is pinned by tests/test_oracle.py (reference known answers,
brute force, invariants)."""

import random

import pytest

pytestmark = pytest.mark.gpu

# N values that exercise segment edges (30 * 144 KiB = 4,423,680 integers per
# default segment; 30 * 128 KiB = 3,932,160 for the smaller tuned shapes), the
# p*p rule of low segments and partial last bytes.
EDGE_N = [7, 8, 10, 29, 30, 31, 100, 101, 1000, 30 * 4096 - 1, 30 * 4096, 3932159, 3932160, 3932161,
          3932160 * 2 + 17, 4423679, 4423680, 4423681, 4423680 * 2 + 17, 10**6, 10**7, 12345677, 10**8]


@pytest.mark.parametrize("N", EDGE_N + [2, 3, 4, 5, 6, 10**9])
def test_prime_count(voss, oracle, N):
    assert voss.primes.prime_count(N) == oracle.prime_count(N)


@pytest.mark.parametrize("N", [10, 100, 10**6, 3932161, 10**7, 10**8, 10**9])
def test_full_stats_vs_oracle(voss, oracle, N):
    cp = 10**7 if N >= 10**8 else 0
    g = voss.stats.full_stats(N, cp_interval=cp)
    o = oracle.full_stats(N, cp_interval=cp)
    for k in ("prime_count", "pi_4_1", "pi_4_3", "pi_3_1", "pi_3_2", "last_prime", "max_gap"):
        assert g[k] == o[k], k
    assert g["hist"] == o["hist"]
    assert g["records"] == o["records"]
    assert g["large_gaps"] == o["large_gaps"]
    assert g["checkpoints"] == o["checkpoints"]


def test_full_stats_random_vs_oracle(voss, oracle):
    # seeded N anywhere in [1e5, 3e9] with random checkpoint intervals and
    # large-gap thresholds: every output against the oracle
    rng = random.Random(2024)
    for _ in range(10):
        N = rng.randrange(10**5, 3 * 10**9)
        cp = rng.choice([0, rng.randrange(10**7, max(N, 10**7 + 1))])
        thr = rng.choice([500, rng.randrange(250, 500)])
        g = voss.stats.full_stats(N, cp_interval=cp, large_threshold=thr)
        o = oracle.full_stats(N, cp_interval=cp, large_threshold=thr)
        for k in ("prime_count", "pi_4_1", "pi_4_3", "pi_3_1", "pi_3_2", "last_prime", "max_gap"):
            assert g[k] == o[k], (N, cp, thr, k)
        assert g["hist"] == o["hist"], (N, cp, thr)
        assert g["records"] == o["records"], (N, cp, thr)
        assert g["large_gaps"] == o["large_gaps"], (N, cp, thr)
        assert g["checkpoints"] == o["checkpoints"], (N, cp, thr)


@pytest.mark.parametrize("N", [2_300_942_869, 2_300_942_900, 2_301_000_000])
def test_record_in_the_last_segments_vs_oracle(voss, oracle, N):
    # the maximal gap 320 ends at 2,300,942,869 (A002386/A005250), in the
    # last, partial segment of the sweep: the records rescan and the
    # checkpoint split run on a short final segment
    g = voss.stats.full_stats(N, cp_interval=10**8, large_threshold=300)
    o = oracle.full_stats(N, cp_interval=10**8, large_threshold=300)
    for k in ("prime_count", "pi_4_1", "pi_4_3", "pi_3_1", "pi_3_2", "last_prime", "max_gap"):
        assert g[k] == o[k], k
    assert g["hist"] == o["hist"]
    assert g["records"] == o["records"]
    assert g["large_gaps"] == o["large_gaps"]
    assert g["checkpoints"] == o["checkpoints"]
    assert g["records"][-1] == (2_300_942_869, 320)


def test_ranges_near_the_top_vs_oracle(voss, oracle):
    # queries between 1e13 and the API's 1e14 ceiling: sieving primes up to
    # 1e7, most of them in the sparse tier of a short segment
    rng = random.Random(99)
    qs = []
    for _ in range(24):
        a = rng.randrange(10**13, 10**14 - 10**7)
        qs.append((a, a + int(10 ** rng.uniform(1, 6.5)) - 1))
    qs += [(10**14 - 10**6, 10**14), (10**14 - 100, 10**14), (10**14 - 30, 10**14 - 29)]
    got = voss.stats.batch_ranges([q[0] for q in qs], [q[1] for q in qs])
    for (a, b), g in zip(qs, got):
        o = oracle.range_stats(a, b)
        assert (g["count"], g["first"], g["last"]) == (o["count"], o["first"], o["last"]), (a, b)
        assert g["hist"] == o["hist"], (a, b)
        assert voss.stats.range_stats(a, b) == g, (a, b)
    assert voss.primes.in_range(10**14 - 10**5, 10**14) == oracle.primes_in_range(10**14 - 10**5, 10**14)


def test_full_stats_numpy_output(voss):
    # the numpy form (histogram buffer handed over uninitialised) equals the
    # list form entry for entry, the zeros past the largest gap included
    import numpy as np
    g = voss.stats.full_stats(10**8, cp_interval=10**7)
    for _ in range(2):
        a = voss.stats.full_stats(10**8, cp_interval=10**7, as_numpy=True)
        assert isinstance(a["hist"], np.ndarray) and len(a["hist"]) == len(g["hist"]) == 100000
        assert a["hist"].tolist() == g["hist"]
        assert [tuple(r) for r in a["checkpoints"].tolist()] == g["checkpoints"]
    assert g["hist"][g["max_gap"] + 1:] == [0] * (100000 - g["max_gap"] - 1)


def test_full_stats_large_gap_threshold(voss, oracle):
    # a low threshold makes the large-gap list long, exercising the append path
    g = voss.stats.full_stats(10**8, large_threshold=150)
    o = oracle.full_stats(10**8, large_threshold=150)
    assert g["large_gaps"] == o["large_gaps"]
    assert g["n_large"] == len(o["large_gaps"])


def test_context_known_answers(voss):
    with voss.primes.Context(10**9) as c:
        assert c.prime_count() == 50_847_534
        assert c.twins() == 3_424_506
        assert c.cousin() == 3_424_679
        assert c.sexy() == 6_089_791
        cb = c.chebyshev()
        assert (cb.pi_4_1, cb.pi_4_3, cb.difference) == (25_423_491, 25_424_042, 551)
        assert c.large_gaps() == []
        s = c.statistics()
        assert s.total_gaps == 50_847_533
        assert round(s.mean_gap, 4) == 19.6666
    with voss.primes.Context(10**6) as c:
        assert (c.twins(), c.cousin(), c.sexy()) == (8169, 8143, 13549)
        assert c.statistics().total_gaps == 78497
        assert c.chebyshev().pi_4_1 == 39175
    with voss.primes.Context(10**7) as c:
        s = c.statistics()
        assert (round(s.mean_gap, 4), round(s.std_dev, 4), round(s.skewness, 4)) == (15.0471, 12.5697, 1.8147)


def test_context_minimal_and_progress(voss):
    seen = []
    with voss.primes.Context(10**9, profile="minimal") as c:
        c.set_progress(lambda *a: seen.append(a))
        assert c.prime_count() == 50_847_534
    assert seen and seen[-1][1] == len(seen) and seen[-1][3] == 50_847_534
    assert sum(a[2] for a in seen) == 50_847_534 - 3


def test_statistics_match_oracle_moments(voss, oracle):
    import math
    o = oracle.full_stats(10**8)
    h = o["hist"]
    tot = sum(h[1:])
    mean = sum(g * h[g] for g in range(1, len(h))) / tot
    var = sum(h[g] * (g - mean) ** 2 for g in range(1, len(h)) if h[g]) / tot
    with voss.primes.Context(10**8) as c:
        s = c.statistics()
    assert s.total_gaps == tot
    assert abs(s.mean_gap - mean) < 1e-12 * mean
    assert abs(s.std_dev - math.sqrt(var)) < 1e-9


def test_export_csv(voss, oracle, tmp_path):
    # library/tests/test_primes.py:490-519 (three files, stats header, total_gaps
    # line) plus every line's content against the oracle (io/csv.cpp:54-110 formats)
    o = oracle.full_stats(10**6)
    prefix = str(tmp_path / "voss_test")
    with voss.primes.Context(10**6) as c:
        files = c.export_csv(prefix)
        s = c.statistics()
    assert files == [prefix + "_chebyshev.csv", prefix + "_large_gaps.csv", prefix + "_stats.csv"]
    cheb = open(files[0]).read().splitlines()
    assert cheb == ["pi_4_1,pi_4_3,difference", f"{o['pi_4_1']},{o['pi_4_3']},{o['pi_4_3'] - o['pi_4_1']}"]
    assert cheb[1] == "39175,39322,147"
    assert open(files[1]).read().splitlines() == ["position,gap"]
    stats = open(files[2]).read().splitlines()
    assert stats[0] == "metric,value" and "total_gaps,78497" in stats
    assert stats[1:5] == [f"mean_gap,{s.mean_gap:.6f}", f"std_dev,{s.std_dev:.6f}", f"skewness,{s.skewness:.6f}",
                          f"kurtosis,{s.kurtosis:.6f}"]
    assert stats[1] == "mean_gap,12.739098"  # (p_last - 2) / (pi - 1) = 999981 / 78497 at 1e6


def test_export_csv_large_gaps_1e12(voss, tmp_path):
    # the 11 gaps >= 500 below 1e12 (library/tests/test_primes.py:447-458), in
    # the list's order: gap descending, then position ascending
    prefix = str(tmp_path / "big")
    with voss.primes.Context(10**12) as c:
        c.export_csv(prefix)
        gaps = c.large_gaps()
    rows = open(prefix + "_large_gaps.csv").read().splitlines()
    assert rows[0] == "position,gap" and len(rows) == 12
    assert rows[1] == "738832928467,540"
    assert rows[1:] == [f"{g.position},{g.gap}" for g in gaps]
    parsed = [tuple(map(int, r.split(","))) for r in rows[1:]]
    assert parsed == sorted(parsed, key=lambda t: (-t[1], t[0]))


@pytest.mark.parametrize("a,b", [(2, 10), (10, 30), (100, 130), (10, 100), (0, 1), (2, 2), (90, 110), (2, 1000),
                                 (10**9, 10**9 + 10**4), (999_999_000, 1_000_001_000), (7, 7), (8, 10),
                                 (3932150, 3932170)])
def test_in_range(voss, oracle, a, b):
    got = voss.primes.in_range(a, b)
    want = oracle.primes_in_range(max(a, 2), b)
    if b > 100:
        want = [p for p in want if p >= 7]  # the reference wheel path omits 2, 3, 5
    assert got == want


def test_in_range_limit(voss):
    with pytest.raises(voss.VossError):
        voss.primes.in_range_with_limit(2, 10**6, 100)
    assert len(voss.primes.in_range_with_limit(1000, 2000, 1000)) == 135


def _rand_queries(rng, n, lo_exp, hi_exp, wlo, whi):
    qs = []
    for _ in range(n):
        a = int(10 ** rng.uniform(lo_exp, hi_exp))
        w = int(10 ** rng.uniform(wlo, whi))
        qs.append((a, a + w - 1))
    return qs


def test_batch_ranges_vs_oracle(voss, oracle):
    rng = random.Random(7)
    qs = _rand_queries(rng, 60, 0, 11, 0, 7.3)
    qs += [(1, 1), (2, 3), (2, 5), (3, 7), (5, 5), (0, 0), (7, 30), (24, 28), (10**10, 10**10), (4, 4), (38, 44), (31, 89), (30, 119), (60, 61), (89, 97), (119, 150)]
    got = voss.stats.batch_ranges([q[0] for q in qs], [q[1] for q in qs])
    for (a, b), g in zip(qs, got):
        o = oracle.range_stats(a, b)
        assert (g["count"], g["first"], g["last"]) == (o["count"], o["first"], o["last"]), (a, b)
        assert g["hist"] == o["hist"], (a, b)


def test_batch_equals_plain(voss):
    rng = random.Random(11)
    qs = _rand_queries(rng, 40, 6, 12, 3, 8)
    batch = voss.stats.batch_ranges([q[0] for q in qs], [q[1] for q in qs])
    for q, bres in zip(qs, batch):
        assert voss.stats.range_stats(*q) == bres


@pytest.mark.parametrize("n", [1, 2, 3, 6, 7, 10, 100, 1000, 10**5, 10**6, 10**7, 5 * 10**7])
def test_nth(voss, oracle, n):
    assert voss.primes.nth(n) == oracle.nth_prime(n)


def test_nth_beyond_reference_cap(voss):
    assert voss.stats.nth_prime(10**8) == 2_038_074_743
    assert voss.stats.nth_prime(10**9) == 22_801_763_489
    with pytest.raises(voss.VossError):
        voss.primes.nth(50_000_001)


def test_goldbach(voss, oracle):
    ns = [4, 6, 8, 10, 100, 1000, 10**4, 10**5, 123456, 10**6, 999998]
    assert [voss.primes.goldbach_count(n) for n in ns] == [oracle.goldbach_count_sieve(n) for n in ns]
    assert voss.stats.goldbach_counts(ns) == oracle.goldbach_counts(ns)
    assert voss.primes.goldbach_partitions(10) == [(3, 7), (5, 5)]
    assert voss.primes.goldbach_partitions(20) == [(3, 17), (7, 13)]
    assert voss.primes.goldbach_partitions(4) == [(2, 2)]
    r = voss.conjectures.test_goldbach_upto(10**7)
    assert r["status"] == "no_counterexample_found_upto_N"


def test_goldbach_counts_many(voss, oracle):
    # every small even n (all residues mod 30, partners across byte borrows),
    # random n up to 1e8 and the edges of a 1e8 bitmap, in one launch
    import random
    rng = random.Random(5)
    ns = list(range(4, 2000, 2)) + [2 * rng.randrange(2, 5 * 10**7) for _ in range(40)] + [10**8, 10**8 - 2]
    assert voss.stats.goldbach_counts(ns) == oracle.goldbach_counts(ns)


def test_sophie_germain(voss, oracle):
    for N, want in [(10, 3), (30, 6), (1000, 37), (10**6, 7746)]:
        with voss.primes.Context(N) as c:
            assert c.sophie_germain() == want
    with voss.primes.Context(10**7 + 3) as c:
        assert c.sophie_germain() == oracle.sophie_germain(10**7 + 3)


def test_mr_batch(voss, oracle):
    rng = random.Random(3)
    xs = [0, 1, 2, 3, 4, 5, 2047, 3215031751, 3825123056546413051, 2**64 - 59, 2**64 - 1, 1000000007,
          18446744073709551557, 4294967297, 4294967291]
    xs += [rng.getrandbits(64) | 1 for _ in range(500)] + [rng.getrandbits(40) for _ in range(500)]
    got = voss.stats.is_prime_batch(xs)
    assert got == [oracle.is_prime(x) for x in xs]


def test_mr_batch_pseudoprimes_and_dense_primes(voss, oracle):
    # strong pseudoprimes to base 2, Carmichael numbers, spsp to bases 2..17
    # and 2..23, the 2^32 edge, and runs made only of primes (every lane of a
    # wave survives base 2, so the shared second stage takes several rounds)
    spsp2 = [2047, 3277, 4033, 4681, 8321, 15841, 29341, 42799, 49141, 52633, 65281, 74665, 80581, 85489,
             88357, 90751, 1194649, 12327121]
    carmichael = [561, 1105, 1729, 2465, 2821, 6601, 8911, 41041, 62745, 63973, 75361, 101101, 126217,
                  172081, 188461, 252601, 278545, 294409, 314821, 334153, 340561, 399001, 410041, 449065]
    big = [341550071728321, 3825123056546413051, 4294967291, 4294967293, 4294967295, 4294967296,
           4294967297, 4294967311, 2**63 - 25, 2**63 + 29, 2**64 - 59, 2**64 - 95]
    rng = random.Random(11)
    primes = []
    for _ in range(200):
        rc, q = oracle.next_prime(rng.getrandbits(63) + 2**63 - 2**62)
        if rc == 0:
            primes.append(q)
    small_primes = [oracle.next_prime(rng.getrandbits(31))[1] for _ in range(100)]
    semis = [a * b for a, b in zip(small_primes[:50], small_primes[50:])]
    xs = spsp2 + carmichael + big + primes + small_primes + semis
    got = voss.stats.is_prime_batch(xs)
    assert got == [oracle.is_prime(x) for x in xs]
    assert all(got[len(spsp2) + len(carmichael) + len(big):][:len(primes) + len(small_primes)])


def test_next_prev_batch(voss, oracle):
    rng = random.Random(5)
    xs = [0, 1, 2, 3, 4, 10, 100, 10**9, 10**12] + [rng.getrandbits(rng.randint(2, 62)) for _ in range(300)]
    nxt = voss.stats.next_prime_batch(xs)
    assert nxt == [oracle.next_prime(x)[1] for x in xs]
    prv = voss.stats.prev_prime_batch(xs)
    assert prv == [(oracle.prev_prime(x)[1] if x > 2 else 0) for x in xs]


@pytest.mark.parametrize("cfg", [(64 * 1024, 1024, 2, 1024), (96 * 1024, 512, 1, 512), (144 * 1024, 1024, 1, 2048)])
def test_configs_agree(voss, oracle, cfg):
    old = voss.stats.get_config()
    try:
        voss.stats.set_config(*cfg)
        o = oracle.full_stats(2 * 10**8 + 7, cp_interval=10**7)
        g = voss.stats.full_stats(2 * 10**8 + 7, cp_interval=10**7)
        for k in ("prime_count", "pi_4_1", "pi_3_1", "hist", "records", "checkpoints"):
            assert g[k] == o[k], (cfg, k)
        assert voss.primes.prime_count(10**9 + 7) == 50847534 + 1  # 10^9+7 is prime
    finally:
        voss.stats.set_config(old["seg_bytes"], old["block"], old["wgs_per_cu"], old["tier1_p"])


def test_conjectures_at_their_limits(voss):
    c = voss.conjectures
    assert c.test_bertrand_upto(10**7)["status"] == "no_counterexample_found_upto_N"
    assert c.test_legendre_upto(10**14)["status"] == "no_counterexample_found_upto_N"
