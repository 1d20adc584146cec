/* test_capi.c — the voss C ABI driven from plain C, as a C caller of the
 * reference library would (library/tests/test_prime_count.cpp:1-77, extended
 * to the Context, range, nth/next/prev, is_prime, Goldbach and factorisation
 * entry points).  Built against include/voss/ and libvoss.so only.
 *
 *   test_capi args   argument checks and error strings: no device call
 *   test_capi full   known answers through every entry point (needs the GPU)
 *
 * Expected values are the reference's own test values (test_primes.py,
 * test_prime_count.cpp) and published tables (OEIS A006880, A007508).
 * Prints one line per check; exits 0 when all pass.
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "voss/voss.h"
#include "voss/voss_primes.h"

static int failures = 0;

static void expect_u64(const char* name, uint64_t got, uint64_t want) {
    if (got == want) {
        printf("  [OK]   %-48s %llu\n", name, (unsigned long long)got);
    } else {
        printf("  [FAIL] %-48s got=%llu want=%llu (last error: %s)\n", name, (unsigned long long)got,
               (unsigned long long)want, voss_get_last_error());
        ++failures;
    }
}

static void expect_int(const char* name, int got, int want) { expect_u64(name, (uint64_t)got, (uint64_t)want); }

static void expect_true(const char* name, int cond) { expect_u64(name, (uint64_t)(cond != 0), 1); }

static uint64_t pi_of(uint64_t n) {
    uint64_t out = 0;
    int rc = voss_primes_prime_count(n, &out);
    if (rc != VOSS_OK) {
        printf("  [FAIL] prime_count(%llu) rc=%d (%s)\n", (unsigned long long)n, rc, voss_get_last_error());
        ++failures;
    }
    return out;
}

static void args_mode(void) {
    printf("Error handling (no device):\n");
    expect_int("strerror(VOSS_OK)[0] == 'O'", voss_strerror(VOSS_OK)[0], 'O');
    expect_true("strerror of every code is non-empty",
                voss_strerror(VOSS_ERR_INVALID_N)[0] && voss_strerror(VOSS_ERR_INTERNAL)[0]);
    {
        uint64_t out = 999;
        expect_int("prime_count(0) returns INVALID_N", voss_primes_prime_count(0, &out), VOSS_ERR_INVALID_N);
        expect_u64("prime_count(0) leaves out untouched", out, 999);
        expect_true("last error is set", voss_get_last_error()[0] != 0);
    }
    expect_int("prime_count(10, NULL) returns INVALID_ARG", voss_primes_prime_count(10, NULL), VOSS_ERR_INVALID_ARG);
    {
        voss_primes_ctx* ctx = (voss_primes_ctx*)0x1;
        expect_int("ctx_new(1) returns INVALID_N", voss_primes_ctx_new(1, VOSS_PROFILE_STANDARD, &ctx),
                   VOSS_ERR_INVALID_N);
        expect_int("ctx_new(profile 9) returns INVALID_ARG", voss_primes_ctx_new(100, 9, &ctx),
                   VOSS_ERR_INVALID_ARG);
        expect_int("ctx_new(out NULL) returns INVALID_ARG", voss_primes_ctx_new(100, 0, NULL), VOSS_ERR_INVALID_ARG);
    }
    {
        uint64_t* arr = (uint64_t*)0x1;
        uint64_t n = 77;
        expect_int("in_range(out NULL) returns INVALID_ARG", voss_primes_in_range(2, 10, NULL, &n),
                   VOSS_ERR_INVALID_ARG);
        expect_int("in_range(b > 1e14) returns OUT_OF_RANGE",
                   voss_primes_in_range(2, 100000000000001ull, &arr, &n), VOSS_ERR_OUT_OF_RANGE);
    }
    {
        uint64_t out = 5;
        expect_int("nth(0) returns INVALID_N", voss_primes_nth(0, &out), VOSS_ERR_INVALID_N);
        expect_int("prev(2) returns INVALID_N", voss_primes_prev(2, &out), VOSS_ERR_INVALID_N);
        expect_int("goldbach_count(7) returns INVALID_N", voss_primes_goldbach_count(7, &out), VOSS_ERR_INVALID_N);
        expect_u64("out untouched by the failed calls", out, 5);
    }
    voss_free(NULL);
    voss_primes_ctx_free(NULL);
    expect_true("voss_free(NULL) and ctx_free(NULL) return", 1);
}

static void full_mode(void) {
    printf("Small values (hand-verified):\n");
    expect_u64("pi(2)", pi_of(2), 1);
    expect_u64("pi(3)", pi_of(3), 2);
    expect_u64("pi(5)", pi_of(5), 3);
    expect_u64("pi(10)", pi_of(10), 4);
    expect_u64("pi(100)", pi_of(100), 25);
    printf("Medium values (OEIS A006880):\n");
    expect_u64("pi(10^6)", pi_of(1000000ull), 78498ull);
    expect_u64("pi(10^7)", pi_of(10000000ull), 664579ull);
    expect_u64("pi(10^8)", pi_of(100000000ull), 5761455ull);
    expect_u64("pi(10^9)", pi_of(1000000000ull), 50847534ull);

    printf("Context at 10^9:\n");
    {
        voss_primes_ctx* ctx = NULL;
        expect_int("ctx_new", voss_primes_ctx_new(1000000000ull, VOSS_PROFILE_STANDARD, &ctx), VOSS_OK);
        uint64_t v = 0, c1 = 0, c3 = 0, nl = 99;
        voss_primes_stats_t s;
        memset(&s, 0, sizeof(s));
        voss_primes_ctx_prime_count(ctx, &v);
        expect_u64("prime_count", v, 50847534ull);
        voss_primes_ctx_twins(ctx, &v);
        expect_u64("twins", v, 3424506ull);
        voss_primes_ctx_cousin(ctx, &v);
        expect_u64("cousin", v, 3424679ull);
        voss_primes_ctx_sexy(ctx, &v);
        expect_u64("sexy", v, 6089791ull);
        voss_primes_ctx_chebyshev_bias(ctx, &c1, &c3);
        expect_u64("pi(x;4,1)", c1, 25423491ull);
        expect_u64("pi(x;4,3)", c3, 25424042ull);
        expect_int("statistics", voss_primes_ctx_statistics(ctx, &s), VOSS_OK);
        expect_u64("total_gaps", s.total_gaps, 50847533ull);
        expect_true("mean gap 19.6666 (4 decimals)", s.mean_gap > 19.66655 && s.mean_gap < 19.66665);
        voss_primes_ctx_large_gaps_count(ctx, &nl);
        expect_u64("no gap >= 500 below 10^9", nl, 0);
        voss_primes_ctx_free(ctx);
    }
    printf("Context at 10^12 (large gaps):\n");
    {
        voss_primes_ctx* ctx = NULL;
        uint64_t nl = 0, pos = 0;
        uint32_t gap = 0;
        expect_int("ctx_new", voss_primes_ctx_new(1000000000000ull, VOSS_PROFILE_FULL, &ctx), VOSS_OK);
        voss_primes_ctx_large_gaps_count(ctx, &nl);
        expect_u64("gaps >= 500 below 10^12", nl, 11);
        voss_primes_ctx_large_gaps_get(ctx, 0, &pos, &gap);
        expect_u64("largest gap", gap, 540);
        expect_u64("its position", pos, 738832928467ull);
        expect_int("get(11) returns INVALID_ARG", voss_primes_ctx_large_gaps_get(ctx, 11, &pos, &gap),
                   VOSS_ERR_INVALID_ARG);
        voss_primes_ctx_free(ctx);
    }
    printf("Minimal profile:\n");
    {
        voss_primes_ctx* ctx = NULL;
        uint64_t v = 0;
        voss_primes_ctx_new(1000000ull, VOSS_PROFILE_MINIMAL, &ctx);
        voss_primes_ctx_prime_count(ctx, &v);
        expect_u64("prime_count", v, 78498ull);
        expect_true("twins rejected", voss_primes_ctx_twins(ctx, &v) != VOSS_OK);
        voss_primes_ctx_free(ctx);
    }
    printf("Ranges:\n");
    {
        uint64_t* arr = NULL;
        uint64_t n = 0;
        expect_int("in_range(1e9, 1e9 + 1e4)", voss_primes_in_range(1000000000ull, 1000010000ull, &arr, &n), VOSS_OK);
        expect_u64("count", n, 487);
        expect_u64("first", n ? arr[0] : 0, 1000000007ull);
        expect_u64("last", n ? arr[n - 1] : 0, 1000009999ull);
        voss_free(arr);
        arr = NULL;
        voss_primes_in_range(2, 10, &arr, &n);
        expect_true("in_range(2, 10) = 2 3 5 7", n == 4 && arr[0] == 2 && arr[1] == 3 && arr[2] == 5 && arr[3] == 7);
        voss_free(arr);
        arr = (uint64_t*)0x1;
        voss_primes_in_range(24, 28, &arr, &n);
        expect_true("empty range gives NULL, 0", arr == NULL && n == 0);
        expect_int("more primes than max_count is OUT_OF_MEMORY",
                   voss_primes_in_range_with_limit(2, 1000000, 100, &arr, &n), VOSS_ERR_OUT_OF_MEMORY);
    }
    printf("nth / next / prev / is_prime:\n");
    {
        static const uint64_t ns[] = {1, 2, 3, 10, 100, 1000, 1000000};
        static const uint64_t ps[] = {2, 3, 5, 29, 541, 7919, 15485863};
        uint64_t v = 0;
        int f = -1;
        for (int i = 0; i < 7; ++i) {
            char name[64];
            snprintf(name, sizeof(name), "nth(%llu)", (unsigned long long)ns[i]);
            voss_primes_nth(ns[i], &v);
            expect_u64(name, v, ps[i]);
        }
        voss_primes_next(1000000000ull, &v);
        expect_u64("next(1e9)", v, 1000000007ull);
        voss_primes_next(1000000000000ull, &v);
        expect_u64("next(1e12)", v, 1000000000039ull);
        voss_primes_prev(1000000000ull, &v);
        expect_u64("prev(1e9)", v, 999999937ull);
        voss_primes_prev(1000000000000ull, &v);
        expect_u64("prev(1e12)", v, 999999999989ull);
        voss_primes_is_prime(1000000007ull, &f);
        expect_int("is_prime(1000000007)", f, 1);
        voss_primes_is_prime(3215031751ull, &f);
        expect_int("is_prime(3215031751) (spsp to bases 2..7)", f, 0);
    }
    printf("Goldbach / factorisation:\n");
    {
        uint64_t v = 0, n = 0;
        uint64_t* arr = NULL;
        voss_primes_goldbach_count(100, &v);
        expect_u64("goldbach_count(100)", v, 6);
        voss_primes_goldbach_count(100000, &v);
        expect_u64("goldbach_count(1e5)", v, 810);
        voss_primes_goldbach_partitions(10, &arr, &n);
        expect_true("partitions(10) = (3,7) (5,5)", n == 2 && arr[0] == 3 && arr[1] == 7 && arr[2] == 5 && arr[3] == 5);
        voss_free(arr);
        arr = (uint64_t*)0x1;
        expect_int("goldbach_test_upto(1e6)", voss_primes_goldbach_test_upto(1000000ull, &arr, &n, 10), VOSS_OK);
        expect_true("no counterexample below 1e6 (NULL, 0)", arr == NULL && n == 0);
        voss_primes_factorize(12000036ull, &arr, &n);
        expect_true("factorize(2^2 * 3 * 1000003)", n == 4 && arr[0] == 2 && arr[1] == 2 && arr[2] == 3 && arr[3] == 1000003);
        voss_free(arr);
    }
}

int main(int argc, char** argv) {
    const int full = argc > 1 && strcmp(argv[1], "full") == 0;
    printf("=== voss C ABI test (%s) ===\n", full ? "full" : "args");
    args_mode();
    if (full) full_mode();
    if (failures == 0) {
        printf("ALL TESTS PASSED\n");
        return 0;
    }
    printf("%d TEST(S) FAILED\n", failures);
    return 1;
}
