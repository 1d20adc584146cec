"""Goldbach counterexample collection on the GPU against the oracle's planted
defect (partners restricted to p <= pmax, which has counterexamples a brute
force lists; the oracle's own form is pinned against one in
tests/test_oracle.py): count, ascending order and the first-max_ce cut, with
the first pass's partner list cut short (scan_limit) so that its candidates
go through the exact-count settling step.  Also the Sophie Germain counts
that the reference's STATE.md publishes."""

import ctypes

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,pmax,cap", [(2000, 13, 10), (3 * 10**6, 31, 40), (3 * 10**6, 7, 10**6),
                                        (3 * 10**6 + 1, 31, 5), (10**5, 3, 7)])
@pytest.mark.parametrize("scan", [0, 3, 5])
def test_planted_defect_lists(voss, oracle, N, pmax, cap, scan):
    want_total, want = oracle.goldbach_test_upto(N, max_ce=cap, pmax=pmax)
    assert want_total > 0
    total, got = voss.stats.goldbach_test_upto(N, max_ce=cap, pmax=pmax, scan_limit=scan)
    assert total == want_total and got == want, (N, pmax, cap, scan)


@pytest.mark.parametrize("N,scan", [(2 * 10**5, 7), (2 * 10**5, 3), (10**6, 13), (10**6 + 1, 0)])
def test_candidates_settled_by_exact_counts(voss, oracle, N, scan):
    # with a short first pass most n are candidates; exact counts over every
    # partner clear them all, as Goldbach holds here
    assert oracle.goldbach_test_upto(N)[0] == 0
    assert voss.stats.goldbach_test_upto(N, max_ce=10, scan_limit=scan) == (0, [])


def test_reference_entry_point(voss, oracle):
    from voss_amd import _capi
    L = _capi._lib
    arr, n = ctypes.POINTER(ctypes.c_uint64)(), ctypes.c_uint64()
    for N in (4, 5, 6, 100, 10**7):
        assert L.voss_primes_goldbach_test_upto(N, ctypes.byref(arr), ctypes.byref(n), 10) == 0
        assert n.value == 0 and not arr
    assert L.voss_primes_goldbach_test_upto(10**12 + 2, ctypes.byref(arr), ctypes.byref(n), 10) == \
        _capi.VOSS_ERR_OUT_OF_RANGE


def test_sophie_germain_published(voss):
    # library/STATE.md:93-94 and library/tests/test_primes.py:561-578
    for N, want in ((10, 3), (30, 6), (100, 10), (1000, 37), (10**6, 7746), (10**8, 423_140),
                    (10**9, 3_308_859)):
        with voss.primes.Context(N) as c:
            assert c.sophie_germain() == want, N
