"""ctypes wrapper over the CPU oracle (liboracle.so) and the reference's own
CPU-only sources compiled into _ref/libvossref_cpu.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker.  The product package
(voss_amd) never imports this module.
"""

import ctypes
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
MAX_GAP = 100000


class OrcGap(ctypes.Structure):
    _fields_ = [("position", ctypes.c_uint64), ("gap", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class OrcStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "prime_count", "pi_4_1", "pi_4_3", "pi_3_1", "pi_3_2", "last_prime",
        "max_gap", "n_large", "n_records", "hist_overflow")]


_lib = None
_ref = None


def _stale(so: Path) -> bool:
    srcs = [_HERE / "voss_oracle.c", _HERE / "voss_oracle.h"]
    return not so.exists() or any(s.stat().st_mtime > so.stat().st_mtime for s in srcs)


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref/ when the reference tree exists)."""
    import subprocess
    so = _HERE / "liboracle.so"
    if force or _stale(so):
        subprocess.run(["make", "-C", str(_HERE), "-B", "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    if Path("/root/reference/library").exists():
        subprocess.run(["make", "-C", str(_HERE), "ref"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        so = _HERE / "liboracle.so"
        if os.environ.get("VOSS_ORACLE_LIB"):  # e.g. the sanitizer build (make -C oracle san)
            so = Path(os.environ["VOSS_ORACLE_LIB"])
        elif _stale(so):
            build()
        L = ctypes.CDLL(str(so))
        u64, u32, i32 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
        P64 = ctypes.POINTER(ctypes.c_uint64)
        L.orc_full_stats.argtypes = [u64, u64, u32, i32, ctypes.POINTER(OrcStats), P64, P64,
                                     ctypes.POINTER(OrcGap), u64, ctypes.POINTER(OrcGap), u64]
        L.orc_full_stats.restype = i32
        L.orc_prime_count.argtypes = [u64, i32]
        L.orc_prime_count.restype = u64
        L.orc_range_stats.argtypes = [u64, u64, i32, P64, P64, P64, P64]
        L.orc_range_stats.restype = i32
        L.orc_primes_in_range.argtypes = [u64, u64, P64, u64]
        L.orc_primes_in_range.restype = u64
        L.orc_is_prime_mr.argtypes = [u64]
        L.orc_is_prime_mr.restype = i32
        for f in ("orc_nth_prime", "orc_next_prime", "orc_prev_prime"):
            getattr(L, f).argtypes = [u64, P64]
            getattr(L, f).restype = i32
        L.orc_goldbach_count_mr.argtypes = [u64]
        L.orc_goldbach_count_mr.restype = u64
        L.orc_goldbach_count_sieve.argtypes = [u64]
        L.orc_goldbach_count_sieve.restype = u64
        L.orc_goldbach_counts.argtypes = [P64, u64, P64, i32]
        L.orc_goldbach_counts.restype = i32
        L.orc_goldbach_test_upto.argtypes = [u64, P64, u64, i32]
        L.orc_goldbach_test_upto.restype = u64
        L.orc_goldbach_test_upto_pmax.argtypes = [u64, u64, P64, u64, i32]
        L.orc_goldbach_test_upto_pmax.restype = u64
        L.orc_sophie_germain.argtypes = [u64]
        L.orc_sophie_germain.restype = u64
        _lib = L
    return _lib


def ref_cpu():
    """The reference's own CPU C++ (built from its sources into _ref/), or None."""
    global _ref
    if _ref is None:
        so = _HERE / "_ref" / "libvossref_cpu.so"
        if not so.exists():
            return None
        _ref = ctypes.CDLL(str(so))
    return _ref


# ---------------------------------------------------------------- wrappers
def full_stats(N, cp_interval=0, large_threshold=500, threads=0, rec_cap=4096, large_cap=1 << 16):
    """Return a dict with every statistic of [2, N] (see voss_oracle.h)."""
    L = lib()
    st = OrcStats()
    hist = (ctypes.c_uint64 * MAX_GAP)()
    ncp = N // cp_interval if cp_interval else 0
    cp = (ctypes.c_uint64 * max(1, 5 * ncp))()
    lg = (OrcGap * large_cap)()
    rec = (OrcGap * rec_cap)()
    rc = L.orc_full_stats(N, cp_interval, large_threshold, threads, ctypes.byref(st), hist,
                          cp if cp_interval else None, lg, large_cap, rec, rec_cap)
    if rc != 0:
        raise ValueError(f"orc_full_stats rc={rc}")
    nl = min(st.n_large, large_cap)
    nr = min(st.n_records, rec_cap)
    return {
        "N": N,
        "prime_count": st.prime_count,
        "pi_4_1": st.pi_4_1, "pi_4_3": st.pi_4_3,
        "pi_3_1": st.pi_3_1, "pi_3_2": st.pi_3_2,
        "last_prime": st.last_prime, "max_gap": st.max_gap,
        "hist": list(hist), "hist_overflow": st.hist_overflow,
        "large_gaps": [(lg[i].position, lg[i].gap) for i in range(nl)],
        "records": [(rec[i].position, rec[i].gap) for i in range(nr)],
        "checkpoints": [tuple(cp[5 * j + f] for f in range(5)) for j in range(ncp)],
    }


def prime_count(N, threads=0):
    return lib().orc_prime_count(N, threads)


def range_stats(a, b, threads=0):
    L = lib()
    c, f, l = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    h = (ctypes.c_uint64 * 257)()
    L.orc_range_stats(a, b, threads, ctypes.byref(c), ctypes.byref(f), ctypes.byref(l), h)
    return {"count": c.value, "first": f.value, "last": l.value, "hist": list(h)}


def primes_in_range(a, b, cap=None):
    L = lib()
    if cap is None:
        cap = L.orc_primes_in_range(a, b, None, 0)
    buf = (ctypes.c_uint64 * max(1, cap))()
    n = L.orc_primes_in_range(a, b, buf, cap)
    return list(buf[:min(n, cap)])


def is_prime(n):
    return bool(lib().orc_is_prime_mr(n))


def _u64_call(fname, x):
    out = ctypes.c_uint64()
    rc = getattr(lib(), fname)(x, ctypes.byref(out))
    return rc, out.value


def nth_prime(n):
    rc, v = _u64_call("orc_nth_prime", n)
    if rc:
        raise ValueError(f"nth_prime({n}) rc={rc}")
    return v


def next_prime(x):
    return _u64_call("orc_next_prime", x)


def prev_prime(x):
    return _u64_call("orc_prev_prime", x)


def goldbach_count_mr(n):
    return lib().orc_goldbach_count_mr(n)


def goldbach_count_sieve(n):
    return lib().orc_goldbach_count_sieve(n)


def goldbach_counts(ns, threads=0):
    k = len(ns)
    arr = (ctypes.c_uint64 * max(1, k))(*ns)
    out = (ctypes.c_uint64 * max(1, k))()
    lib().orc_goldbach_counts(arr, k, out, threads)
    return list(out[:k])


def goldbach_test_upto(N, max_ce=10, threads=0, pmax=None):
    """(number of even n in [4, N] with no partition, the first max_ce of them).
    pmax restricts partners to p <= pmax (a planted defect for pinning)."""
    buf = (ctypes.c_uint64 * max(1, max_ce))()
    if pmax is None:
        n = lib().orc_goldbach_test_upto(N, buf, max_ce, threads)
    else:
        n = lib().orc_goldbach_test_upto_pmax(N, pmax, buf, max_ce, threads)
    return n, list(buf[:min(n, max_ce)])


def sophie_germain(N):
    return lib().orc_sophie_germain(N)


def reference_large_gaps(large_gaps, N):
    """The large-gap list the reference's Context would report for N: the
    reference's histogram stitch adds segment-boundary gaps to the histogram
    only (context.cu:472-478), so a gap whose primes fall in different
    reference segments (SEG_NUM = min(ceil(N/30)*30+10, 5e10+10),
    context.cu:260-269) is absent from its list."""
    max_seg = 50000000010
    seg = ((N + 29) // 30) * 30 + 10 if N < max_seg else max_seg
    return [(pos, g) for (pos, g) in large_gaps if (pos - g - 1) // seg == (pos - 1) // seg]
