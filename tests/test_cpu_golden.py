"""The committed batched-range fixture agrees with the oracle on a sample
(the full fixture took minutes of CPU to write)."""

import json
import random
import sys
from pathlib import Path

GOLD = Path(__file__).parent / "golden" / "ranges_4096.json"
sys.path.insert(0, str(GOLD.parent))
from make_ranges_4096 import hist_digest, queries  # noqa: E402


def test_fixture_queries_and_sample(oracle):
    data = json.loads(GOLD.read_text())
    gold = data["queries"]
    assert [(g[0], g[1]) for g in gold] == queries()
    rng = random.Random(0)
    small = [g for g in gold if g[1] - g[0] < 3 * 10**6]
    for g in rng.sample(small, 24):
        r = oracle.range_stats(g[0], g[1])
        assert [r["count"], r["first"], r["last"], hist_digest(r["hist"])] == g[2:]
