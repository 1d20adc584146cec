"""Autotuner for the engine's launch configuration.

Times candidate configurations (segment bytes x block size x workgroups per
CU x tier-1 prime bound) on two representative workloads and writes the
fastest of each to tuned.cfg next to libvoss.so, which the engine reads at
start-up:

  sweep  full statistics over [2, N] (default N = 1e11, the Context workload)
  batch  a sample of the batched-range workload (log-uniform a in
         [1e6, 1e13], width in [1e3, 1e9]; default 4096 queries, seed 0)

The search is coordinate-wise: the segment/block/occupancy shape first (LDS
permitting: segment LDS x workgroups per CU <= 160 KiB), then the tier-1
bound for the best shape.  Every candidate's result is checked against the
first one bit for bit, so a configuration can only win by being faster.

    python -m voss_amd.autotune [--n 1e11] [--queries 4096] [--out PATH]
"""

import argparse
import json
import time
from pathlib import Path

from . import stats

SHAPES = [  # (seg_bytes, block, wgs_per_cu)
    (147456, 1024, 1), (131072, 1024, 1), (114688, 1024, 1), (98304, 1024, 1),
    (65536, 1024, 2), (73728, 512, 1), (65536, 512, 2), (147456, 512, 1),
]
TIER1 = [768, 1024, 1536, 2048, 3072]


def _lds_ok(seg, wgs):
    return (10688 + seg + 16) * wgs <= 160 * 1024  # the kernels keep 10,688 B of tables next to the segment


def _time(fn, reps):
    fn()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        stats.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def _queries(n, seed=0):
    import numpy as np
    rng = np.random.default_rng(seed)
    a = np.floor(10 ** rng.uniform(6, 13, n)).astype(np.uint64)
    w = np.floor(10 ** rng.uniform(3, 9, n)).astype(np.uint64)
    return [int(x) for x in a], [int(x + y - 1) for x, y in zip(a, w)]


def tune(kind, N=10**11, nq=1024, reps=2, log=print):
    if kind == "sweep":
        def run():
            r = stats.full_stats(N, cp_interval=10**9, as_numpy=True)
            return (r["prime_count"], r["pi_4_1"], r["pi_3_1"], int(r["hist"].sum()), r["records"][-1])
    else:
        qa, qb = _queries(nq)

        def run():
            res = stats.batch_ranges(qa, qb, as_numpy=True)
            return (res["count"].tobytes(), res["first"].tobytes(), res["last"].tobytes(), res["hist"].tobytes())
    ref = None
    results = []

    def trial(seg, blk, wgs, t1):
        nonlocal ref
        try:
            stats.set_config(seg, blk, wgs, t1, kind=kind)
        except Exception as e:  # invalid for this device
            log(f"  skip {seg} {blk} {wgs} {t1}: {e}")
            return None
        out = run()
        if ref is None:
            ref = out
        elif out != ref:
            raise RuntimeError(f"configuration {(seg, blk, wgs, t1)} changed the {kind} result")
        t = _time(run, reps)
        results.append({"seg_bytes": seg, "block": blk, "wgs_per_cu": wgs, "tier1_p": t1, "ms": t * 1e3})
        log(f"  {kind} seg={seg} block={blk} wgs={wgs} tier1={t1}: {t * 1e3:.3f} ms")
        return t

    best = None
    for seg, blk, wgs in SHAPES:
        if not _lds_ok(seg, wgs):
            continue
        t = trial(seg, blk, wgs, 2048)
        if t is not None and (best is None or t < best[0]):
            best = (t, seg, blk, wgs, 2048)
    _, seg, blk, wgs, _ = best
    for t1 in TIER1:
        if t1 == 2048:
            continue
        t = trial(seg, blk, wgs, t1)
        if t is not None and t < best[0]:
            best = (t, seg, blk, wgs, t1)
    t, seg, blk, wgs, t1 = best
    stats.set_config(seg, blk, wgs, t1, kind=kind)
    return {"seg_bytes": seg, "block": blk, "wgs_per_cu": wgs, "tier1_p": t1, "ms": t * 1e3}, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=float, default=1e11)
    ap.add_argument("--queries", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=None, help="tuned.cfg path (default: next to libvoss.so)")
    ap.add_argument("--log", default=None, help="also write every trial as JSON here")
    args = ap.parse_args()
    from . import _capi
    out = Path(args.out) if args.out else Path(_capi.LIBRARY_PATH).with_name("tuned.cfg")
    best_s, trials_s = tune("sweep", N=int(args.n), reps=args.reps)
    best_b, trials_b = tune("batch", nq=args.queries, reps=args.reps)
    dev = stats.device_info()
    out.write_text(f"sweep {best_s['seg_bytes']} {best_s['block']} {best_s['wgs_per_cu']} {best_s['tier1_p']}\n"
                   f"batch {best_b['seg_bytes']} {best_b['block']} {best_b['wgs_per_cu']} {best_b['tier1_p']}\n")
    summary = {"device": dev, "sweep": best_s, "batch": best_b, "trials": {"sweep": trials_s, "batch": trials_b}}
    if args.log:
        Path(args.log).write_text(json.dumps(summary, indent=1))
    print(json.dumps({"device": dev, "sweep": best_s, "batch": best_b, "written": str(out)}))


if __name__ == "__main__":
    main()
