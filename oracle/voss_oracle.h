/*
 * voss_oracle.h — CPU restatement of the voss wheel-30 sieve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against, and the "port" CPU baseline timed by
 * bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product library (libvoss.so) never links
 * or calls it.
 *
 * Conventions follow the reference exactly (file:line into the reference):
 *   - bitmap layout: byte k bit o <-> 30k + W30[o],
 *     W30 = {1,7,11,13,17,19,23,29}            (src/voss_w30_v7.cu:38-39)
 *   - 2, 3, 5 are seeds outside the bitmap; count starts at 3,
 *     hist[1] += 1 (2->3), hist[2] += 1 (3->5), the 5->7 gap is an
 *     ordinary boundary gap               (library/src/primes/context.cu:321-322,513-518)
 *   - pi(x;4,1) includes 5, pi(x;4,3) includes 3, 2 in neither
 *                                          (context.cu:520-524, chebyshev.cu:11-25)
 *   - N < 7 short-circuits with pi = 1/2/3 and hist[1]/hist[2] set, and
 *     the mod-4 tallies left at 0           (context.cu:228-247)
 *   - large gaps: gap >= threshold (500), position = the larger prime,
 *     sorted by gap descending              (gaps.cu:55-66, context.cu:546-550)
 * Additions with no reference counterpart (stated conventions):
 *   - pi(x;3,1) / pi(x;3,2): 3 in neither class, 2 counts in class 2, so
 *     c31 + c32 + 1 = pi(x) for x >= 3.
 *   - maximal-gap records: gaps strictly larger than every earlier gap,
 *     starting with (3, gap 1); position = larger prime.
 */
#ifndef VOSS_ORACLE_H
#define VOSS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_GAP 100000 /* histogram length, as VOSS_MAX_GAP (context.cu:63) */

typedef struct {
    uint64_t position; /* larger prime of the gap */
    uint32_t gap;
    uint32_t pad;
} orc_gap_t;

typedef struct {
    uint64_t prime_count;
    uint64_t pi_4_1, pi_4_3;     /* reference convention (context.cu:520-524) */
    uint64_t pi_3_1, pi_3_2;     /* stated convention above */
    uint64_t last_prime;
    uint64_t max_gap;
    uint64_t n_large;            /* number of gaps >= threshold found (may exceed cap) */
    uint64_t n_records;          /* number of records (may exceed cap) */
    uint64_t hist_overflow;      /* gaps >= ORC_MAX_GAP (never below 2^64 in practice) */
} orc_stats_t;

/* Full statistics over [2, N].
 * hist: ORC_MAX_GAP u64 bins (bin g = number of consecutive-prime gaps equal to g).
 * cp_interval: if nonzero, checkpoints at x = j*cp_interval (j = 1..N/cp_interval);
 *   cp_out receives 5 u64 per checkpoint: pi, pi_4_1, pi_4_3, pi_3_1, pi_3_2.
 * large/large_cap: gaps >= large_threshold, sorted (gap desc, position asc).
 * records/rec_cap: maximal gaps in increasing position order.
 * threads: worker threads (0 = all online cores).
 * Returns 0 on success. */
int orc_full_stats(uint64_t N, uint64_t cp_interval, uint32_t large_threshold,
                   int threads, orc_stats_t* out, uint64_t* hist,
                   uint64_t* cp_out, orc_gap_t* large, uint64_t large_cap,
                   orc_gap_t* records, uint64_t rec_cap);

/* Count-only pi(N) over [2, N] (same sieve, no statistics). */
uint64_t orc_prime_count(uint64_t N, int threads);

/* Range statistics over primes p with a <= p <= b (true primes, 2/3/5 included).
 * hist257[i] for i = 1..256 counts gaps equal to 2i; hist257[0] counts every
 * other gap (1 or > 512).  first/last are 0 when the range holds no prime. */
int orc_range_stats(uint64_t a, uint64_t b, int threads, uint64_t* count,
                    uint64_t* first, uint64_t* last, uint64_t* hist257);

/* All primes in [a, b] ascending (true primes).  Writes at most cap values;
 * returns the total number found. */
uint64_t orc_primes_in_range(uint64_t a, uint64_t b, uint64_t* out, uint64_t cap);

/* Reference restatements of the CPU-only entry points. */
int orc_is_prime_mr(uint64_t n);                       /* miller_rabin.cpp:26-51 */
int orc_nth_prime(uint64_t n, uint64_t* out);          /* nth.cu:40-103 semantics (no 5e7 cap) */
int orc_next_prime(uint64_t x, uint64_t* out);         /* next_prev.cu:22-49 */
int orc_prev_prime(uint64_t x, uint64_t* out);         /* next_prev.cu:51-83 */
uint64_t orc_goldbach_count_mr(uint64_t n);            /* goldbach.cpp:21-43 (MR loop) */
uint64_t orc_goldbach_count_sieve(uint64_t n);         /* same count from a sieve */
/* Goldbach partition counts for every n in ns (even, >= 4, <= nmax) from one sieve. */
int orc_goldbach_counts(const uint64_t* ns, uint64_t k, uint64_t* out, int threads);
/* Even n in [4, N] with no partition, ascending, at most max_ce (goldbach_upto.cu:25-46). */
uint64_t orc_goldbach_test_upto(uint64_t N, uint64_t* ce, uint64_t max_ce, int threads);
uint64_t orc_goldbach_test_upto_pmax(uint64_t N, uint64_t pmax, uint64_t* ce, uint64_t max_ce, int threads);
/* Sophie Germain primes p <= N (p and 2p+1 prime) (sophie_germain_upto.cu:22-33). */
uint64_t orc_sophie_germain(uint64_t N);

#ifdef __cplusplus
}
#endif

#endif /* VOSS_ORACLE_H */
