"""The large sweeps: full statistics over [2, 1e11] against the oracle run on
the same box, [2, 1e12] against the committed oracle fixture (bins 2..1500,
mod-4/mod-3 tallies every 1e10, records, gaps >= 500), published known
answers, and the Goldbach check to 1e9 with partition counts at multiples
of 1e6."""

import json
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

FIX = Path(__file__).parent / "golden" / "sweep_1e12.json"


def test_full_stats_1e11_vs_oracle(voss, oracle):
    N = 10**11
    g = voss.stats.full_stats(N, cp_interval=10**9)
    o = oracle.full_stats(N, cp_interval=10**9)
    for k in ("prime_count", "pi_4_1", "pi_4_3", "pi_3_1", "pi_3_2", "last_prime", "max_gap"):
        assert g[k] == o[k], k
    assert g["hist"] == o["hist"]
    assert g["records"] == o["records"]
    assert g["large_gaps"] == o["large_gaps"]
    assert g["checkpoints"] == o["checkpoints"]
    # published values: pi(1e11) (A006880), twin count (A007508), the last
    # maximal gap below 1e11 (464 after 42,652,618,343; A002386/A005250)
    assert g["prime_count"] == 4_118_054_813
    assert g["hist"][2] == 224_376_048
    assert g["records"][-1] == (42_652_618_343 + 464, 464)


def test_context_progress_splits_1e11(voss, oracle):
    # the progress callback fires once per reference segment of SEG_NUM =
    # min(ceil(N/30)*30 + 10, 5e10 + 10) integers (context.cu:260-269,
    # 433-436): at 1e11 two calls, (seg, 2, primes in seg, cumulative), the
    # first segment being (0, 5e10 + 10].  Every Context result must equal
    # the callback-free sweep's and the published values.
    N = 10**11
    S = 50_000_000_010

    def results(c):
        # the result classes compare by identity, as the reference's do
        cb, st = c.chebyshev(), c.statistics()
        return (c.prime_count(), c.twins(), c.cousin(), c.sexy(), c.sophie_germain(), cb.pi_4_1, cb.pi_4_3,
                st.mean_gap, st.std_dev, st.skewness, st.kurtosis, st.total_gaps, c.large_gaps())

    seen = []
    with voss.primes.Context(N, profile="full") as c:
        c.set_progress(lambda *a: seen.append(a))
        split = results(c)
    with voss.primes.Context(N, profile="full") as c:
        whole = results(c)
    pi_s = oracle.prime_count(S)
    assert seen == [(1, 2, pi_s - 3, pi_s), (2, 2, 4_118_054_813 - pi_s, 4_118_054_813)]
    assert split == whole
    assert split[0] == 4_118_054_813 and split[1] == 224_376_048


@pytest.mark.parametrize("N", [5 * 10**10, 50_000_000_010, 50_000_000_011, 10**11 + 7, 150_000_000_031])
def test_context_progress_segments(voss, oracle, N):
    # one call below 5e10 + 10 (N = 5e10: SEG_NUM = 5e10 + 20, NUM_SEG = 1);
    # N = 5e10 + 11 has a second segment of one integer, inside the kernel
    # segment that the first launch already sieved
    S = ((N + 29) // 30) * 30 + 10 if N < 50_000_000_010 else 50_000_000_010
    nseg = (N + S - 1) // S
    bounds = [min((j + 1) * S, N) for j in range(nseg)]
    pis = [oracle.prime_count(b) for b in bounds]
    want, prev = [], 3
    for j, pi in enumerate(pis):
        want.append((j + 1, nseg, pi - prev, pi))
        prev = pi
    for profile in ("minimal", "standard"):
        seen = []
        with voss.primes.Context(N, profile=profile) as c:
            c.set_progress(lambda *a: seen.append(a))
            assert c.prime_count() == pis[-1]
        assert seen == want, (profile, seen, want)


def _fixture():
    if not FIX.exists():
        pytest.skip("tests/golden/sweep_1e12.json not written")
    return json.loads(FIX.read_text())


def test_sweep_1e12_vs_fixture(voss):
    f = _fixture()
    N = f["N"]
    g = voss.stats.full_stats(N, cp_interval=f["checkpoint_interval"])
    for k in ("prime_count", "pi_4_1", "pi_4_3", "pi_3_1", "pi_3_2", "last_prime", "max_gap"):
        assert g[k] == f[k], k
    hist = {str(i): c for i, c in enumerate(g["hist"][:1501]) if c}
    assert hist == f["hist_1_1500"]
    assert [list(r) for r in g["records"]] == f["records"]
    assert [list(r) for r in g["large_gaps"]] == f["large_gaps_ge_500"]
    assert [list(c) for c in g["checkpoints"]] == f["checkpoints"]
    # published: pi(1e12) = 37,607,912,018; 11 gaps >= 500, the largest 540
    # ending at 738,832,928,467 (library/tests/test_primes.py:447-458)
    assert g["prime_count"] == 37_607_912_018
    assert len(g["large_gaps"]) == 11 and g["large_gaps"][0] == (738_832_928_467, 540)


def test_goldbach_1e9(voss):
    f = _fixture()
    r = voss.conjectures.test_goldbach_upto(10**9)
    assert r["status"] == "no_counterexample_found_upto_N"
    assert f["goldbach_counterexamples_upto_1e9"] == 0
    ns = [j * 10**6 for j in range(1, 1001)]
    assert voss.stats.goldbach_counts(ns) == f["goldbach_counts"]


# P(6)/P(2), the ratio of consecutive-prime gaps 6 and 2, as the frozen v7
# program records its trend (src/voss_w30_v7.cu:524-534, 4 decimals)
V7_P6_P2 = {10**9: 1.7783, 10**10: 1.8018, 10**11: 1.8208, 10**12: 1.8366, 10**13: 1.8497}


@pytest.mark.parametrize("N", [10**9, 10**10, 10**11, 10**12])
def test_v7_gap_repulsion_trend(voss, N):
    h = voss.stats.full_stats(N)["hist"]
    assert abs(h[6] / h[2] - V7_P6_P2[N]) < 5e-5


def test_prime_count_1e13(voss):
    # library/STATE.md:102-103: pi(1e13) = 346,065,536,839
    assert voss.primes.prime_count(10**13) == 346_065_536_839


def test_full_stats_1e13_known(voss):
    # published: pi(1e13) (A006880), twin pairs below 1e13 (A007508) and the
    # last maximal gap below 1e13, 674 after 7,177,162,611,713 (A002386)
    g = voss.stats.full_stats(10**13, cp_interval=10**12)
    assert g["prime_count"] == 346_065_536_839
    assert g["hist"][2] == 15_834_664_872
    assert g["records"][-1] == (7_177_162_611_713 + 674, 674)
    assert g["max_gap"] == 674
    assert abs(g["hist"][6] / g["hist"][2] - V7_P6_P2[10**13]) < 5e-5
    h = g["hist"]
    assert sum(h) == g["prime_count"] - 1 and sum(i * c for i, c in enumerate(h)) == g["last_prime"] - 2
    assert g["pi_4_1"] + g["pi_4_3"] + 1 == g["prime_count"] == g["pi_3_1"] + g["pi_3_2"] + 1
    assert g["checkpoints"][-1][0] == g["prime_count"]


def test_prime_count_1e14(voss):
    # the API's upper limit; pi(1e14) = 3,204,941,750,802 (A006880)
    assert voss.primes.prime_count(10**14) == 3_204_941_750_802
