"""The C ABI without a GPU: libvoss.so loads, exports every function the
headers in include/voss declare, the Python binding covers them all, argument
checks fire before any device work, and device entry points fail loudly
(VOSS_ERR_CUDA, the reference's code for "no device") instead of falling back
to the CPU."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / "include" / "voss").glob("*.h"))
DECL = re.compile(r"^[A-Za-z_][\w\s\*]*?\b(voss_\w+)\s*\(", re.M)


def declared():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith(("#", "typedef")))
        names.update(DECL.findall(text))
    return names


def test_headers_declare_the_reference_surface():
    names = declared()
    assert len(names) >= 60
    for must in ("voss_primes_prime_count", "voss_primes_ctx_new", "voss_primes_goldbach_test_upto",
                 "voss_primes_nth", "voss_numbers_phi", "voss_amd_full_stats", "voss_amd_batch_ranges",
                 "voss_amd_server_replay"):
        assert must in names


# The reference library's 38 exported entry points (library/include/voss/
# voss.h:19-25, voss_primes.h:13-181, voss_numbers.h:16-45), the set its CI
# checks with `nm -D`: every one must be declared here and exported.
REFERENCE_EXPORTS = (
    ["voss_strerror", "voss_get_last_error", "voss_free"]
    + ["voss_primes_" + n for n in (
        "prime_count", "in_range", "in_range_with_limit", "nth", "next", "prev", "is_prime",
        "factorize", "is_mersenne_prime", "is_fermat_prime",
        "goldbach_count", "goldbach_partitions", "goldbach_test_upto")]
    + ["voss_primes_ctx_" + n for n in (
        "new", "free", "prime_count", "twins", "cousin", "sexy", "statistics", "chebyshev_bias",
        "large_gaps_count", "large_gaps_get", "export_csv", "set_progress", "sophie_germain",
        "sophie_germain_upto")]
    + ["voss_numbers_" + n for n in ("phi", "tau", "sigma", "mu", "gcd", "lcm", "fibonacci", "factorial")]
)


def test_reference_exports_declared_and_exported():
    assert len(REFERENCE_EXPORTS) == 38 == len(set(REFERENCE_EXPORTS))
    names = declared()
    assert [n for n in REFERENCE_EXPORTS if n not in names] == []
    lib = ctypes.CDLL(str(ROOT / "voss_amd" / "libvoss.so"))
    assert [n for n in REFERENCE_EXPORTS if not hasattr(lib, n)] == []


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(ROOT / "voss_amd" / "libvoss.so"))
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_symbol(voss):
    from voss_amd import _capi
    assert declared() <= set(_capi._SIGNATURES), sorted(declared() - set(_capi._SIGNATURES))


def _no_device(voss):
    try:
        voss.stats.device_info()
    except voss.VossError:
        return True
    return False


def test_argument_checks_need_no_device(voss):
    from voss_amd import _capi
    L = _capi._lib
    out = ctypes.c_uint64()
    assert L.voss_primes_prime_count(1, ctypes.byref(out)) == _capi.VOSS_ERR_INVALID_N
    assert L.voss_primes_prime_count(10**14 + 1, ctypes.byref(out)) == _capi.VOSS_ERR_OUT_OF_RANGE
    assert L.voss_primes_nth(0, ctypes.byref(out)) == _capi.VOSS_ERR_INVALID_N
    assert L.voss_primes_goldbach_count(7, ctypes.byref(out)) == _capi.VOSS_ERR_INVALID_N
    assert L.voss_amd_nth_prime(10**12 + 1, ctypes.byref(out)) == _capi.VOSS_ERR_OUT_OF_RANGE
    assert L.voss_amd_server_submit(None, ctypes.byref(out)) == _capi.VOSS_ERR_INVALID_ARG
    # host-only entry points work without a device
    assert L.voss_numbers_phi(36, ctypes.byref(out)) == 0 and out.value == 12
    assert L.voss_primes_next(100, ctypes.byref(out)) == 0 and out.value == 101
    cfg = _capi.AmdServerConfig()
    assert L.voss_amd_server_default_config(ctypes.byref(cfg)) == 0
    assert (cfg.slots, cfg.max_batch, cfg.nth_max, cfg.goldbach_max) == (4, 1024, 10**12, 10**8)


_NO_DEVICE_SCRIPT = r"""
import ctypes, sys
sys.path.insert(0, sys.argv[1])
import voss_amd as voss
from voss_amd import _capi
L = _capi._lib
out = ctypes.c_uint64()
E = _capi.VOSS_ERR_CUDA
# the reference maps "no device" to VOSS_ERR_CUDA (prime_count.cu:147-151, 316-318)
assert L.voss_primes_prime_count(10**9, ctypes.byref(out)) == E
assert b"no HIP device" in L.voss_get_last_error()
ctx = ctypes.c_void_p()
assert L.voss_primes_ctx_new(10**6, 1, ctypes.byref(ctx)) == 0
assert L.voss_primes_ctx_prime_count(ctx, ctypes.byref(out)) == E
L.voss_primes_ctx_free(ctx)
arr, n = ctypes.POINTER(ctypes.c_uint64)(), ctypes.c_uint64()
assert L.voss_primes_in_range(1000, 2000, ctypes.byref(arr), ctypes.byref(n)) == E
assert L.voss_primes_goldbach_test_upto(1000, ctypes.byref(arr), ctypes.byref(n), 10) == E
assert L.voss_amd_server_start(None) == E
# host-only entry points still answer
assert L.voss_primes_prime_count(5, ctypes.byref(out)) == 0 and out.value == 3
assert L.voss_primes_next(100, ctypes.byref(out)) == 0 and out.value == 101
try:
    voss.primes.prime_count(10**6)
    raise SystemExit("no error raised")
except voss.VossError:
    pass
print("ok")
"""


def test_device_paths_fail_loudly_without_a_device():
    # a child process with the devices hidden, so this holds on a GPU box too
    import os
    import subprocess
    import sys
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, "-c", _NO_DEVICE_SCRIPT, str(ROOT)], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr


def test_serving_trace_is_deterministic(voss):
    from voss_amd import serve
    r1, a1 = serve.make_trace(4000, rate=2000.0, seed=1)
    r2, a2 = serve.make_trace(4000, rate=2000.0, seed=1)
    assert np.array_equal(r1, r2) and np.array_equal(a1, a2)
    share = np.bincount(r1["kind"], minlength=6) / len(r1)
    assert abs(share[serve.RANGE] - 0.5) < 0.03
    assert abs(share[[serve.NTH, serve.NEXT, serve.PREV]].sum() - 0.2) < 0.03
    assert abs(share[serve.IS_PRIME] - 0.2) < 0.03
    assert abs(share[serve.GOLDBACH] - 0.1) < 0.03
    assert abs(len(a1) / a1[-1] - 2000.0) < 150
    g = r1["a"][r1["kind"] == serve.GOLDBACH]
    assert (g % 2 == 0).all() and g.min() >= 4 and g.max() <= 10**8
    rq = r1[r1["kind"] == serve.RANGE]
    w = rq["b"] - rq["a"] + 1
    assert rq["a"].min() >= 10**6 and rq["a"].max() <= 10**13 and w.min() >= 10**3 and w.max() <= 10**9
    nnp = r1["a"][np.isin(r1["kind"], [serve.NTH, serve.NEXT, serve.PREV])]
    assert nnp.min() >= 1 and nnp.max() <= 10**12
    assert (r1["n"][r1["kind"] == serve.IS_PRIME] == 64).all()


def test_c_program_argument_checks():
    """The plain-C ABI test (tests/c/test_capi.c) in its no-device mode: error
    codes, untouched out-parameters, last-error text, NULL releases."""
    import subprocess
    exe = ROOT / "tests" / "c" / "test_capi"
    if not exe.exists():
        subprocess.run(["make", "-C", str(ROOT / "tests" / "c")], check=True, capture_output=True)
    r = subprocess.run([str(exe), "args"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL TESTS PASSED" in r.stdout
