"""Writes sweep_1e12.json: the oracle's answers for the large single-GPU sweep.

  - full statistics over [2, 1e12]: pi, pi(x;4,1), pi(x;4,3), pi(x;3,1),
    pi(x;3,2), last prime, the gap histogram (every gap 1..1500 that occurs;
    gaps above 1500 are listed in large), maximal-gap records, gaps >= 500,
    and the five tallies at every multiple of 1e10;
  - Goldbach partition counts (p1 <= p2) for n = j * 1e6, j = 1..1000, and the
    number of even n in [4, 1e9] with no partition (expected 0).

    python tests/golden/make_sweep_1e12.py   (tens of minutes on 8 cores)
"""

import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

N = 10**12
CP = 10**10


def main():
    from oracle import oracle as O
    O.build()
    t0 = time.time()
    s = O.full_stats(N, cp_interval=CP, large_threshold=500)
    print(f"full_stats({N:.0e}) {time.time() - t0:.0f} s", flush=True)
    hist = {str(g): c for g, c in enumerate(s["hist"][:1501]) if c}
    out = {
        "N": N, "checkpoint_interval": CP,
        "prime_count": s["prime_count"], "pi_4_1": s["pi_4_1"], "pi_4_3": s["pi_4_3"],
        "pi_3_1": s["pi_3_1"], "pi_3_2": s["pi_3_2"], "last_prime": s["last_prime"], "max_gap": s["max_gap"],
        "hist_1_1500": hist, "records": s["records"], "large_gaps_ge_500": s["large_gaps"],
        "checkpoints": s["checkpoints"],
    }
    t0 = time.time()
    ns = [j * 10**6 for j in range(1, 1001)]
    out["goldbach_ns"] = "j * 1e6, j = 1..1000"
    out["goldbach_counts"] = O.goldbach_counts(ns)
    print(f"goldbach_counts {time.time() - t0:.0f} s", flush=True)
    t0 = time.time()
    nce, _ = O.goldbach_test_upto(10**9, max_ce=10)
    out["goldbach_counterexamples_upto_1e9"] = nce
    print(f"goldbach_test_upto {time.time() - t0:.0f} s", flush=True)
    Path(__file__).with_name("sweep_1e12.json").write_text(json.dumps(out, separators=(",", ":")))


if __name__ == "__main__":
    main()
