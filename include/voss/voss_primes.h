/* voss_primes.h — prime-counting, gap statistics and range queries.
 *
 * Drop-in for the reference header library/include/voss/voss_primes.h:13-181.
 * Every entry point returns a VOSS_* code; out-parameters are written only on
 * success.  Each declaration below names the reference entry it replaces.
 */
#ifndef VOSS_PRIMES_H
#define VOSS_PRIMES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* pi(N) for 2 <= N <= 1e14.  Replaces prime_count.cu:298-322. */
int voss_primes_prime_count(uint64_t N, uint64_t* out);

#define VOSS_PROFILE_MINIMAL   0
#define VOSS_PROFILE_STANDARD  1
#define VOSS_PROFILE_FULL      2

/* Opaque context: N, profile and a lazily computed cache (context.cu:70-98). */
typedef struct voss_primes_ctx voss_primes_ctx;

/* context.cu:566-599 */
int voss_primes_ctx_new(uint64_t N, int profile, voss_primes_ctx** out_ctx);
/* context.cu:601-603 (NULL accepted) */
void voss_primes_ctx_free(voss_primes_ctx* ctx);

/* context.cu:652-667 */
int voss_primes_ctx_prime_count(voss_primes_ctx* ctx, uint64_t* out);
/* Consecutive-prime gaps of 2, 4 and 6 (context.cu:965-975). */
int voss_primes_ctx_twins(voss_primes_ctx* ctx, uint64_t* out);
int voss_primes_ctx_cousin(voss_primes_ctx* ctx, uint64_t* out);
int voss_primes_ctx_sexy(voss_primes_ctx* ctx, uint64_t* out);

/* Primes p <= N with 2p + 1 prime (context.cu:765-894). */
int voss_primes_ctx_sophie_germain(voss_primes_ctx* ctx, uint64_t* out);
int voss_primes_ctx_sophie_germain_upto(voss_primes_ctx* ctx, uint64_t* out);

typedef struct {
    double   mean_gap;
    double   std_dev;
    double   skewness;
    double   kurtosis;    /* excess */
    uint64_t total_gaps;
} voss_primes_stats_t;

/* context.cu:694-755 */
int voss_primes_ctx_statistics(voss_primes_ctx* ctx, voss_primes_stats_t* out);
/* context.cu:940-963 */
int voss_primes_ctx_chebyshev_bias(voss_primes_ctx* ctx, uint64_t* out_pi_4_1, uint64_t* out_pi_4_3);
/* Gaps >= 500, sorted by gap descending (context.cu:896-938). */
int voss_primes_ctx_large_gaps_count(voss_primes_ctx* ctx, uint64_t* out_count);
int voss_primes_ctx_large_gaps_get(voss_primes_ctx* ctx, uint64_t index,
                                   uint64_t* out_position, uint32_t* out_gap);
/* <prefix>_chebyshev.csv, _large_gaps.csv, _stats.csv (io/csv.cpp:17-115). */
int voss_primes_ctx_export_csv(voss_primes_ctx* ctx, const char* prefix);

/* Primes in [a, b] as a malloc'd array freed by voss_free (range.cu:267-323). */
int voss_primes_in_range(uint64_t a, uint64_t b, uint64_t** out_array, uint64_t* out_count);
int voss_primes_in_range_with_limit(uint64_t a, uint64_t b, uint64_t max_count,
                                    uint64_t** out_array, uint64_t* out_count);

/* nth.cu:40-103, next_prev.cu:22-83 */
int voss_primes_nth(uint64_t n, uint64_t* out_prime);
int voss_primes_next(uint64_t x, uint64_t* out_prime);
int voss_primes_prev(uint64_t x, uint64_t* out_prime);

/* Deterministic Miller-Rabin (is_prime.cpp:11-18). */
int voss_primes_is_prime(uint64_t x, int* out);

/* factorize.cpp:84-141, mersenne_fermat.cpp */
int voss_primes_factorize(uint64_t x, uint64_t** out_array, uint64_t* out_count);
int voss_primes_is_mersenne_prime(uint32_t p, int* out);
int voss_primes_is_fermat_prime(uint32_t n, int* out);

/* Called after each internal segment: (seg, total, primes_in_seg, cumulative, user)
 * (voss_primes.h:149-156 of the reference). */
typedef void (*voss_progress_cb)(int seg, int total, uint64_t primes_in_seg,
                                 uint64_t cumulative, void* user);
int voss_primes_ctx_set_progress(voss_primes_ctx* ctx, voss_progress_cb cb, void* user);

/* goldbach.cpp:47-114 */
int voss_primes_goldbach_count(uint64_t n, uint64_t* out);
int voss_primes_goldbach_partitions(uint64_t n, uint64_t** out_array, uint64_t* out_count);
/* goldbach_upto.cu:48-150 */
int voss_primes_goldbach_test_upto(uint64_t N, uint64_t** out_counterexamples,
                                   uint64_t* out_count, uint64_t max_counterexamples);

#ifdef __cplusplus
}
#endif
#endif /* VOSS_PRIMES_H */
