"""Continuous-batching request server (no reference counterpart).

The reference answers each query with its own call (library/src/primes/
prime_count.cu:296-320, nth.cu:40-103, next_prev.cu, core/is_prime.cpp,
goldbach.cpp:21-69).  Here one native scheduler thread (voss_amd/csrc/
serve.cpp) admits mixed requests in arrival order into whichever device
stream is free, within a device-memory budget, and keeps the sieving-prime
table, a prefix-count index for nth and the Goldbach bitmap resident.

    from voss_amd import serve
    reqs, arrivals = serve.make_trace(8000, rate=2000.0, seed=1)
    with serve.Server() as srv:
        responses, stats = srv.replay(reqs, arrivals)

`plain=True` runs the same kernels one request at a time on one stream
(the per-query path) for comparison.
"""

import ctypes

import numpy as np

from . import _capi
from .primes import _ERRMAP

_L = _capi._lib

RANGE, NTH, NEXT, PREV, IS_PRIME, GOLDBACH = range(6)
KIND_NAMES = ("range", "nth", "next", "prev", "is_prime", "goldbach")

REQUEST_DTYPE = np.dtype([("kind", "<u4"), ("n", "<u4"), ("a", "<u8"), ("b", "<u8"), ("x", "<u8", (64,))])
RESPONSE_DTYPE = np.dtype([("status", "<i4"), ("pad", "<u4"), ("value", "<u8"), ("first", "<u8"),
                           ("last", "<u8"), ("hist", "<u4", (257,)), ("pad2", "<u4"),
                           ("latency_ms", "<f8"), ("tick", "<u8")])
assert REQUEST_DTYPE.itemsize == ctypes.sizeof(_capi.AmdRequest)
assert RESPONSE_DTYPE.itemsize == ctypes.sizeof(_capi.AmdResponse)


def make_trace(n: int, rate: float = 2000.0, seed: int = 1, range_share: float = 0.5,
               nnp_share: float = 0.2, is_prime_share: float = 0.2):
    """The mixed serving workload: Poisson arrivals at `rate` per second.

    Kinds: range_share ranges (a = floor(10**U(6, 13)), width floor(10**U(3, 9)),
    as the batched-range workload), nnp_share nth/next/prev in equal parts with
    argument floor(10**U(0, 12)) (nth's n at least 1), is_prime_share batches of
    64 uniform uint64, the rest Goldbach counts of even n uniform in [4, 1e8].
    Drawn from numpy.random.default_rng(seed).  Returns (requests, arrival_s).
    """
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    arrivals -= arrivals[0]
    u = rng.uniform(0.0, 1.0, n)
    reqs = np.zeros(n, dtype=REQUEST_DTYPE)
    c1, c2, c3 = range_share, range_share + nnp_share, range_share + nnp_share + is_prime_share
    is_range = u < c1
    is_nnp = (u >= c1) & (u < c2)
    is_isp = (u >= c2) & (u < c3)
    is_gb = u >= c3
    a = np.floor(10 ** rng.uniform(6, 13, n)).astype(np.uint64)
    w = np.floor(10 ** rng.uniform(3, 9, n)).astype(np.uint64)
    arg = np.maximum(np.floor(10 ** rng.uniform(0, 12, n)), 1).astype(np.uint64)
    which = rng.integers(0, 3, n)
    xs = rng.integers(0, np.iinfo(np.uint64).max, (n, 64), dtype=np.uint64, endpoint=True)
    gb = (2 * rng.integers(2, 50_000_001, n)).astype(np.uint64)
    reqs["kind"][is_range] = RANGE
    reqs["a"][is_range] = a[is_range]
    reqs["b"][is_range] = a[is_range] + w[is_range] - 1
    reqs["kind"][is_nnp] = np.array([NTH, NEXT, PREV], dtype=np.uint32)[which[is_nnp]]
    reqs["a"][is_nnp] = arg[is_nnp]
    reqs["kind"][is_isp] = IS_PRIME
    reqs["n"][is_isp] = 64
    reqs["x"][is_isp] = xs[is_isp]
    reqs["kind"][is_gb] = GOLDBACH
    reqs["a"][is_gb] = gb[is_gb]
    return reqs, arrivals


def request(kind: int, a: int = 0, b: int = 0, x=()) -> np.ndarray:
    """One request record."""
    r = np.zeros((), dtype=REQUEST_DTYPE)
    r["kind"], r["a"], r["b"], r["n"] = kind, a, b, len(x)
    if len(x):
        if len(x) > 64:
            raise ValueError("at most 64 values per is_prime request")
        r["x"][:len(x)] = np.asarray(x, dtype=np.uint64)
    return r


class Server:
    """Process-wide server (start replaces a running one)."""

    def __init__(self, slots: int = None, max_batch: int = None, hbm_reserve: int = None,
                 hbm_budget: int = None, nth_max: int = None, goldbach_max: int = None,
                 plain: bool = False):
        cfg = _capi.AmdServerConfig()
        _capi._check(_L.voss_amd_server_default_config(ctypes.byref(cfg)), _ERRMAP)
        for name, v in (("slots", slots), ("max_batch", max_batch), ("hbm_reserve", hbm_reserve),
                        ("hbm_budget", hbm_budget), ("nth_max", nth_max), ("goldbach_max", goldbach_max)):
            if v is not None:
                setattr(cfg, name, int(v))
        cfg.plain = 1 if plain else 0
        self.config = {name: getattr(cfg, name) for name, _ in cfg._fields_ if name != "pad"}
        _capi._check(_L.voss_amd_server_start(ctypes.byref(cfg)), _ERRMAP)
        self._open = True

    def close(self):
        if self._open:
            self._open = False
            _capi._check(_L.voss_amd_server_stop(), _ERRMAP)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, req) -> int:
        r = np.ascontiguousarray(np.asarray(req, dtype=REQUEST_DTYPE).reshape(()))
        t = ctypes.c_uint64()
        rc = _L.voss_amd_server_submit(ctypes.cast(r.ctypes.data, ctypes.POINTER(_capi.AmdRequest)),
                                       ctypes.byref(t))
        _capi._check(rc, _ERRMAP)
        return t.value

    def wait(self, ticket: int) -> np.ndarray:
        out = np.zeros((), dtype=RESPONSE_DTYPE)
        rc = _L.voss_amd_server_wait(ticket, ctypes.cast(out.ctypes.data, ctypes.POINTER(_capi.AmdResponse)))
        _capi._check(rc, _ERRMAP)
        return out

    def run_one(self, req) -> np.ndarray:
        """Submit one request and wait for it (the request runs alone if nothing else is queued)."""
        return self.wait(self.submit(req))

    def replay(self, reqs: np.ndarray, arrival_s: np.ndarray):
        """Open-loop replay of a trace; returns (responses, stats dict)."""
        reqs = np.ascontiguousarray(reqs, dtype=REQUEST_DTYPE)
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        if len(arr) != len(reqs):
            raise ValueError("one arrival time per request")
        out = np.zeros(len(reqs), dtype=RESPONSE_DTYPE)
        st = _capi.AmdServeStats()
        rc = _L.voss_amd_server_replay(reqs.ctypes.data, arr.ctypes.data, len(reqs), out.ctypes.data,
                                       ctypes.byref(st))
        _capi._check(rc, _ERRMAP)
        return out, {name: getattr(st, name) for name, _ in st._fields_}
