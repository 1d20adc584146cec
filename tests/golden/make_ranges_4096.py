"""Writes ranges_4096.json: the oracle's answers for the batched-range workload.

Workload (seed 0): 4096 queries, a = floor(10**U(6, 13)), w = floor(10**U(3, 9)),
b = a + w - 1, drawn with numpy.random.default_rng(0) (a's first, then w's).
Per query the fixture keeps count, first, last and a 64-bit FNV-1a digest of
the 257-bin gap histogram (bins [i] = gaps of 2i for i = 1..256, [0] = other).

    python tests/golden/make_ranges_4096.py   (several minutes on 8 cores)
"""

import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))


def queries(n=4096, seed=0):
    rng = np.random.default_rng(seed)
    a = np.floor(10 ** rng.uniform(6, 13, n)).astype(np.uint64)
    w = np.floor(10 ** rng.uniform(3, 9, n)).astype(np.uint64)
    return [(int(x), int(x + y - 1)) for x, y in zip(a, w)]


def hist_digest(h):
    d = 0xcbf29ce484222325
    for v in h:
        for k in range(8):
            d ^= (int(v) >> (8 * k)) & 0xFF
            d = (d * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return f"{d:016x}"


def main():
    from oracle import oracle as O
    O.build()
    qs = queries()
    out = []
    for i, (a, b) in enumerate(qs):
        r = O.range_stats(a, b)
        out.append([a, b, r["count"], r["first"], r["last"], hist_digest(r["hist"])])
        if i % 256 == 0:
            print(i, flush=True)
    Path(__file__).with_name("ranges_4096.json").write_text(json.dumps(
        {"workload": "4096 ranges, seed 0, a=10^U(6,13), w=10^U(3,9)", "fields": ["a", "b", "count", "first", "last", "hist_fnv1a64"],
         "queries": out}, separators=(",", ":")))


if __name__ == "__main__":
    main()
