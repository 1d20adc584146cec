"""Batched ranges on the GPU at full size (4096 queries, seed 0) against the
oracle's committed answers (tests/golden/ranges_4096.json, written by
tests/golden/make_ranges_4096.py), and batch == plain per-query path."""

import json
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).parent / "golden" / "ranges_4096.json"
sys.path.insert(0, str(GOLD.parent))
from make_ranges_4096 import hist_digest, queries  # noqa: E402


def test_4096_ranges_batch_matches_golden(voss):
    gold = json.loads(GOLD.read_text())["queries"]
    qs = queries()
    assert [(g[0], g[1]) for g in gold] == qs
    for rep in range(2):  # the second call replays the captured graph
        res = voss.stats.batch_ranges([a for a, _ in qs], [b for _, b in qs])
        for g, r in zip(gold, res):
            assert (r["count"], r["first"], r["last"]) == (g[2], g[3], g[4]), (g[0], g[1], rep)
            assert hist_digest(r["hist"]) == g[5], (g[0], g[1], rep)


def test_4096_ranges_plain_equals_batch(voss):
    qs = queries()[:96]
    batch = voss.stats.batch_ranges([a for a, _ in qs], [b for _, b in qs])
    for q, r in zip(qs, batch):
        assert voss.stats.range_stats(*q) == r


def test_batch_shapes_replay(voss, oracle):
    # different capacity classes and repeated replays with other contents
    import random
    rng = random.Random(9)
    for n in (1, 7, 64, 65, 300):
        qs = []
        for _ in range(n):
            a = rng.randint(0, 10**rng.randint(1, 12))
            qs.append((a, a + rng.randint(0, 10**rng.randint(0, 7))))
        for _ in range(2):
            got = voss.stats.batch_ranges([a for a, _ in qs], [b for _, b in qs])
            for (a, b), g in zip(qs[:20], got[:20]):
                o = oracle.range_stats(a, b)
                assert (g["count"], g["first"], g["last"], g["hist"]) == (o["count"], o["first"], o["last"], o["hist"])


def test_batch_numpy_records_fully_written(voss):
    # the numpy path hands the library an uninitialised array: every field of
    # every record, empty queries (b < a, no wheel byte) and the seeds 2, 3, 5
    # included, must equal the zero-initialised list path; a freed array of
    # the same size full of ones makes stale bytes likely
    import numpy as np
    qs = [(10, 5), (0, 1), (2, 2), (0, 6), (3, 3), (8, 10), (7, 7), (100, 130), (24, 28), (0, 0),
          (10**12, 10**12 + 10**4), (5, 5), (1, 7)]
    for rep in range(3):
        junk = np.full(len(qs) * 1056, 0xFF, dtype=np.uint8)
        del junk
        arr = voss.stats.batch_ranges([a for a, _ in qs], [b for _, b in qs], as_numpy=True)
        lst = voss.stats.batch_ranges([a for a, _ in qs], [b for _, b in qs])
        assert len(arr) == len(lst) == len(qs)
        for q, r, l in zip(qs, arr, lst):
            assert (int(r["count"]), int(r["first"]), int(r["last"])) == (l["count"], l["first"], l["last"]), q
            assert [int(h) for h in r["hist"]] == l["hist"], q
            assert int(r["pad"]) == 0, q
        qs = qs[::-1]
    assert len(voss.stats.batch_ranges([], [], as_numpy=True)) == 0
