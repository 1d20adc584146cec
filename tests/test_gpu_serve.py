"""The request server: each result is the same alone and in the mix, equals
the plain entry points and the oracle, and admission respects the device
memory budget."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_TRACE = 1200


def _same(x, y):
    return (x["status"] == y["status"] and x["value"] == y["value"] and x["first"] == y["first"]
            and x["last"] == y["last"] and np.array_equal(x["hist"], y["hist"]))


@pytest.fixture(scope="module")
def serve(voss):
    from voss_amd import serve as S
    return S


@pytest.fixture(scope="module")
def trace(serve):
    return serve.make_trace(N_TRACE, rate=2000.0, seed=1)


@pytest.fixture(scope="module")
def served(serve, trace):
    reqs, arr = trace
    print(f"serve: replay {len(reqs)} requests at 2000/s", flush=True)
    with serve.Server() as srv:
        mixed, st = srv.replay(reqs, arr)
        burst, st_burst = srv.replay(reqs, np.zeros(len(reqs)))
        print("serve: each request alone", flush=True)
        alone = np.stack([srv.run_one(r) for r in reqs])
    return mixed, st, burst, st_burst, alone


def test_each_result_same_alone_and_in_mix(trace, served):
    reqs, _ = trace
    mixed, st, burst, _, alone = served
    assert (mixed["status"] == 0).all(), np.unique(mixed["status"], return_counts=True)
    for i in range(len(reqs)):
        assert _same(mixed[i], alone[i]), (i, reqs[i]["kind"], reqs[i]["a"], reqs[i]["b"])
        assert _same(burst[i], alone[i]), (i, reqs[i]["kind"], reqs[i]["a"], reqs[i]["b"])
    assert st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]


def test_burst_is_batched(served):
    _, _, burst, st_burst, _ = served
    # everything queued at once: ticks carry many requests and overlap on streams
    assert st_burst["max_tick"] > 1
    assert st_burst["ticks"] < N_TRACE
    assert len(np.unique(burst["tick"])) == st_burst["ticks"]


def test_matches_plain_entry_points(voss, serve, trace, served):
    reqs, _ = trace
    mixed = served[0]
    st = voss.stats
    k = reqs["kind"]
    rng = np.nonzero(k == serve.RANGE)[0]
    got = st.batch_ranges([int(v) for v in reqs["a"][rng]], [int(v) for v in reqs["b"][rng]], as_numpy=True)
    assert np.array_equal(mixed["value"][rng], got["count"])
    assert np.array_equal(mixed["first"][rng], got["first"])
    assert np.array_equal(mixed["last"][rng], got["last"])
    assert np.array_equal(mixed["hist"][rng], got["hist"])
    for kind, fn in ((serve.NEXT, st.next_prime_batch), (serve.PREV, st.prev_prime_batch)):
        idx = np.nonzero(k == kind)[0]
        assert list(mixed["value"][idx]) == fn([int(v) for v in reqs["a"][idx]])
    idx = np.nonzero(k == serve.GOLDBACH)[0]
    assert list(mixed["value"][idx]) == st.goldbach_counts([int(v) for v in reqs["a"][idx]])
    idx = np.nonzero(k == serve.IS_PRIME)[0]
    flat = st.is_prime_batch([int(v) for v in reqs["x"][idx].ravel()])
    for j, i in enumerate(idx):
        bits = flat[64 * j:64 * (j + 1)]
        assert int(mixed["value"][i]) == sum(int(b) << t for t, b in enumerate(bits))
    idx = [i for i in np.nonzero(k == serve.NTH)[0] if reqs["a"][i] <= 10**9]
    assert idx
    for i in idx:
        assert int(mixed["value"][i]) == st.nth_prime(int(reqs["a"][i]))


def test_matches_oracle(oracle, serve, trace, served):
    reqs, _ = trace
    mixed = served[0]
    checked = 0
    for i, r in enumerate(reqs):
        kind, a, b, v = int(r["kind"]), int(r["a"]), int(r["b"]), int(mixed[i]["value"])
        if kind == serve.RANGE and b - a < 2 * 10**6:
            o = oracle.range_stats(a, b)
            assert (v, int(mixed[i]["first"]), int(mixed[i]["last"])) == (o["count"], o["first"], o["last"])
            assert list(mixed[i]["hist"]) == list(o["hist"])
        elif kind == serve.NEXT:
            assert (0, v) == oracle.next_prime(a)
        elif kind == serve.PREV:
            rc, p = oracle.prev_prime(a)
            assert v == (p if rc == 0 else 0)
        elif kind == serve.IS_PRIME:
            assert v == sum(int(oracle.is_prime(int(x))) << t for t, x in enumerate(r["x"]))
        elif kind == serve.GOLDBACH and a <= 2 * 10**6:
            assert v == oracle.goldbach_count_sieve(a)
        elif kind == serve.NTH and a <= 2 * 10**7:
            assert v == oracle.nth_prime(a)
        else:
            continue
        checked += 1
    assert checked > N_TRACE // 3


def test_nth_edges(serve):
    # index boundaries: tiny n, n across the first segment, n = nth_max
    with serve.Server(nth_max=10**6, goldbach_max=10**4) as srv:
        known = {1: 2, 2: 3, 3: 5, 4: 7, 5: 11, 6: 13, 25: 97, 168: 997, 1229: 9973,
                 78498: 999983, 10**6: 15485863}
        for n, p in known.items():
            r = srv.run_one(serve.request(serve.NTH, n))
            assert (int(r["status"]), int(r["value"])) == (0, p), n
        r = srv.run_one(serve.request(serve.NTH, 10**6 + 1))
        assert int(r["status"]) == 2  # beyond nth_max: VOSS_ERR_OUT_OF_RANGE


def test_plain_mode_same_results(serve, trace, served):
    reqs, _ = trace
    alone = served[4]
    n = 300
    with serve.Server(plain=True) as srv:
        out, st = srv.replay(reqs[:n], np.zeros(n))
    assert st["max_tick"] == 1 and st["ticks"] == n
    for i in range(n):
        assert _same(out[i], alone[i]), i


def test_memory_admission(serve, trace, served):
    reqs, _ = trace
    alone = served[4]
    keep = np.nonzero((reqs["kind"] != serve.NTH) | (reqs["a"] <= 10**6))[0][:400]
    budget = 24 * 1024
    with serve.Server(hbm_budget=budget, nth_max=10**6) as srv:
        out, st = srv.replay(reqs[keep], np.zeros(len(keep)))
    assert st["deferred_by_memory"] > 0
    assert st["peak_inflight_bytes"] <= budget
    n_oom = 0
    for j, i in enumerate(keep):
        if int(out[j]["status"]) == 4:  # a range whose working set exceeds the whole budget
            assert int(reqs[i]["kind"]) == serve.RANGE
            n_oom += 1
        else:
            assert _same(out[j], alone[i]), i
    assert n_oom < len(keep) // 2


def test_invalid_requests(serve):
    with serve.Server(nth_max=10**5, goldbach_max=10**5) as srv:
        cases = [(serve.request(serve.NTH, 0), 1), (serve.request(serve.NTH, 10**5 + 1), 2),
                 (serve.request(serve.GOLDBACH, 7), 1), (serve.request(serve.GOLDBACH, 2), 1),
                 (serve.request(serve.GOLDBACH, 10**5 + 2), 2),
                 (serve.request(serve.RANGE, 1, 10**14 + 1), 2),
                 (serve.request(serve.PREV, 2), 0), (serve.request(serve.NEXT, 2**64 - 59), 2)]
        bad = serve.request(serve.IS_PRIME, x=[1, 2])
        bad["n"] = 65
        cases.append((bad, 6))
        bad_kind = serve.request(serve.RANGE, 1, 2)
        bad_kind["kind"] = 9
        cases.append((bad_kind, 6))
        for req, code in cases:
            r = srv.run_one(req)
            assert int(r["status"]) == code, (req, int(r["status"]))
        r = srv.run_one(serve.request(serve.PREV, 2))
        assert int(r["value"]) == 0
        r = srv.run_one(serve.request(serve.RANGE, 2, 1))
        assert (int(r["status"]), int(r["value"])) == (0, 0)


def test_nth_at_scale(voss, serve, served):
    # published values (A006988): p(1e11) and p(1e12); answered from the
    # index a server built, both through the server and the plain entry point
    known = {10**11: 2_760_727_302_517, 10**12: 29_996_224_275_833}
    for n, p in known.items():
        assert voss.stats.nth_prime(n) == p
    with serve.Server() as srv:
        for n, p in known.items():
            r = srv.run_one(serve.request(serve.NTH, n))
            assert (int(r["status"]), int(r["value"])) == (0, p)


def test_ticket_of_an_earlier_server(voss, serve):
    a = serve.Server(nth_max=10**4, goldbach_max=10**4)
    t = a.submit(serve.request(serve.NEXT, 100))
    assert int(a.wait(t)["value"]) == 101
    t2 = a.submit(serve.request(serve.NEXT, 1000))
    a.close()  # answers t2 before it goes away
    with serve.Server(nth_max=10**4, goldbach_max=10**4) as b:
        with pytest.raises(voss.VossError):
            b.wait(t2)
        assert int(b.run_one(serve.request(serve.PREV, 100))["value"]) == 97


def test_cached_ranges(voss, oracle, serve):
    # ranges inside the server's resident bitmap of [7, goldbach_max] are read
    # from it instead of sieved: every edge against the oracle and against the
    # sieving path (stats.range_stats), including b at the cache limit and one
    # past it (sieved again)
    top = 2 * 10**6
    qs = [(2, 10), (2, 1000), (5, 7), (7, 7), (0, 0), (1, 1), (24, 28), (999_990, 1_000_010),
          (top - 100, top), (top - 100, top + 1), (top, top + 50), (30 * 4096 - 5, 30 * 4096 + 5),
          (7, top), (2, top)]
    rng = np.random.default_rng(40)
    for _ in range(40):
        a = int(rng.integers(0, top))
        qs.append((a, min(top, a + int(10 ** rng.uniform(0, 6.3)))))
    with serve.Server(nth_max=10**4, goldbach_max=top) as srv:
        for a, b in qs:
            r = srv.run_one(serve.request(serve.RANGE, a, b))
            o = oracle.range_stats(max(a, 2), b) if b >= 2 else {"count": 0, "first": 0, "last": 0, "hist": [0] * 257}
            got = (int(r["status"]), int(r["value"]), int(r["first"]), int(r["last"]))
            assert got == (0, o["count"], o["first"], o["last"]), (a, b)
            assert list(r["hist"]) == list(o["hist"]), (a, b)
            p = voss.stats.range_stats(a, b)
            assert (p["count"], p["first"], p["last"], p["hist"]) == (got[1], got[2], got[3], list(r["hist"])), (a, b)
