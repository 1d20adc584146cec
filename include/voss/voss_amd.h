/* voss_amd.h — entry points beyond the reference C ABI.
 *
 * These have no counterpart in the reference (library/include/voss/voss.h and voss_primes.h);
 * they expose what the MI355X engine computes in one pass (checkpointed
 * Chebyshev tallies, maximal-gap records, the full histogram), batched and
 * served range queries, GPU Miller-Rabin and Goldbach counts, and the engine
 * configuration.  All return VOSS_* codes and follow the same conventions as
 * voss_primes.h (out-parameters written only on success, thread-local last
 * error).  Implemented by voss_amd/csrc/capi.cpp and serve.cpp.
 */
#ifndef VOSS_AMD_H
#define VOSS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t position; /* larger prime of the gap */
    uint32_t gap;
    uint32_t pad;
} voss_amd_gap_t;

typedef struct {
    uint64_t prime_count;
    uint64_t pi_4_1, pi_4_3;  /* reference convention: 5 in class 1, 3 in class 3, 2 in neither */
    uint64_t pi_3_1, pi_3_2;  /* 3 in neither, 2 in class 2: pi_3_1 + pi_3_2 + 1 = pi for N >= 3 */
    uint64_t last_prime;
    uint64_t max_gap;
    uint64_t n_large;         /* gaps >= large_threshold found */
    uint64_t n_records;       /* maximal-gap records found */
    uint64_t n_checkpoints;   /* N / cp_interval */
} voss_amd_full_stats_t;

/* Full statistics over [2, N] in one call.
 *  hist:        100000 u64 (index = gap; hist[1] = 1 for 2->3), or NULL
 *  checkpoints: 5 u64 per checkpoint x = j*cp_interval, j = 1..N/cp_interval:
 *               pi(x), pi(x;4,1), pi(x;4,3), pi(x;3,1), pi(x;3,2); at most cp_cap
 *               checkpoints written; cp_interval 0 disables; otherwise it must be at
 *               least 30 x the segment bytes (3.9e6 by default)
 *  large:       gaps >= large_threshold sorted (gap desc, position asc), at most large_cap
 *  records:     maximal gaps (each larger than every earlier gap) in ascending
 *               position, starting (3,1), (5,2); at most rec_cap */
int voss_amd_full_stats(uint64_t N, uint64_t cp_interval, uint32_t large_threshold,
                        voss_amd_full_stats_t* out, uint64_t* hist,
                        uint64_t* checkpoints, uint64_t cp_cap,
                        voss_amd_gap_t* large, uint64_t large_cap,
                        voss_amd_gap_t* records, uint64_t rec_cap);

/* Range statistics over the true primes in [a, b] (2, 3, 5 included).
 * hist[i], i = 1..256, counts consecutive-prime gaps equal to 2i; hist[0] counts
 * every other gap (1, or above 512).  first/last are 0 when the range holds no prime. */
typedef struct {
    uint64_t count;
    uint64_t first;
    uint64_t last;
    uint32_t hist[257];
    uint32_t pad;
} voss_amd_range_t;

/* One query, one launch (the plain per-query path). */
int voss_amd_range_stats(uint64_t a, uint64_t b, voss_amd_range_t* out);
/* n queries in one persistent launch over the flattened segment work-list. */
int voss_amd_batch_ranges(const uint64_t* a, const uint64_t* b, uint64_t n, voss_amd_range_t* out);
/* Prime count in [a, b]. */
int voss_amd_count_range(uint64_t a, uint64_t b, uint64_t* out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the count of voss_amd_count_range, written to d_out[0]
 * (device memory) by kernels enqueued on `stream` (a hipStream_t; NULL is the
 * null stream).  It allocates nothing, copies nothing and does not wait: its
 * work counter and per-segment counts live in a fixed 256 KiB device array
 * per GPU, so [a, b] may span at most 16384 segments (7.2e10 integers at the
 * default 144 KiB segments; VOSS_ERR_OUT_OF_RANGE beyond).  The library must
 * already be initialised on the current device by an earlier device call
 * (e.g. voss_amd_device_info); otherwise VOSS_ERR_INVALID_ARG.  Calls that
 * overlap on one GPU race on the array: serialise them (one stream). */
int voss_amd_count_range_noalloc(uint64_t a, uint64_t b, uint64_t* d_out, void* stream);

/* GPU primality of n values (deterministic for every u64); out[i] = 0/1. */
int voss_amd_is_prime_batch(const uint64_t* x, uint64_t n, uint8_t* out);
/* GPU next/prev prime of n values (same semantics as voss_primes_next/prev,
 * except that where those fail the batch writes 0 and still returns VOSS_OK:
 * prev of x <= 2, next of x >= 2^64 - 59). */
int voss_amd_next_prime_batch(const uint64_t* x, uint64_t n, uint64_t* out);
int voss_amd_prev_prime_batch(const uint64_t* x, uint64_t n, uint64_t* out);
/* nth prime without the reference's 5e7 cap (n <= 1e12). */
int voss_amd_nth_prime(uint64_t n, uint64_t* out);
/* Goldbach partition counts (p1 <= p2) for even n >= 4, one launch for all. */
int voss_amd_goldbach_counts(const uint64_t* ns, uint64_t k, uint64_t* out);
/* voss_primes_goldbach_test_upto with two more knobs: partners restricted to
 * primes p <= pmax (0: no restriction; a restriction plants counterexamples:
 * the even n in [6, N] with no p <= min(n/2, pmax) such that n - p is prime),
 * and the first pass's partner list cut at scan_limit (0: every sieving
 * prime), past which its candidates are settled by exact partition counts.
 * Writes the first max_ce counterexamples in ascending order (malloc'd, free
 * with voss_free) and, if out_total is not NULL, how many there are in all.
 * N <= 1e12. */
int voss_amd_goldbach_test_upto_pmax(uint64_t N, uint64_t pmax, uint64_t scan_limit, uint64_t max_ce,
                                     uint64_t** out_ce, uint64_t* out_count, uint64_t* out_total);

/* ---------------------------------------------------------------- serving
 * A continuous-batching request server (one per process).  Requests of mixed
 * kinds are admitted in arrival order into whichever of `slots` device
 * streams is free, as many per tick as max_batch and the device-memory budget
 * allow (working bytes in flight <= min(hbm_budget, free HBM - hbm_reserve)),
 * so work that arrives while other ticks run overlaps on the device.  The
 * sieving-prime table, a prefix-count index for nth (n <= nth_max) and the
 * wheel bitmap of [0, goldbach_max] stay resident.  Each response equals what
 * the plain entry points return for the same request. */
enum {
    VOSS_AMD_REQ_RANGE = 0,    /* a, b: count (value), first, last, hist as voss_amd_range_t */
    VOSS_AMD_REQ_NTH = 1,      /* a = n (1-based): value = p_n */
    VOSS_AMD_REQ_NEXT = 2,     /* a = x: value = smallest prime > x */
    VOSS_AMD_REQ_PREV = 3,     /* a = x: value = largest prime < x, 0 when x <= 2 */
    VOSS_AMD_REQ_IS_PRIME = 4, /* x[0..n): value bit i = x[i] is prime */
    VOSS_AMD_REQ_GOLDBACH = 5  /* a = even n >= 4: value = #{p1 <= p2 : p1 + p2 = n} */
};

typedef struct {
    uint32_t kind;
    uint32_t n;       /* is_prime: values in x (<= 64) */
    uint64_t a, b;
    uint64_t x[64];
} voss_amd_request_t;

typedef struct {
    int32_t status;   /* VOSS_OK or the code the plain entry point would return */
    uint32_t pad;
    uint64_t value;
    uint64_t first, last;
    uint32_t hist[257];
    uint32_t pad2;
    double latency_ms; /* completion minus arrival */
    uint64_t tick;     /* serial number of the batch that carried the request */
} voss_amd_response_t;

typedef struct {
    int32_t slots;         /* concurrent device streams (1..8); default 4 */
    uint32_t max_batch;    /* requests per tick; default 1024 */
    uint64_t hbm_reserve;  /* bytes left free on the device; default 1 GiB */
    uint64_t hbm_budget;   /* working bytes in flight; default 8 GiB */
    uint64_t nth_max;      /* default 10^12 */
    uint64_t goldbach_max; /* default 10^8 */
    int32_t plain;         /* 1: one request per tick on one stream (the per-query path) */
    int32_t pad;
} voss_amd_server_config_t;

typedef struct {
    double wall_s;        /* first arrival to last completion */
    double qps;           /* requests / wall_s */
    double p50_ms, p99_ms, mean_ms, max_ms;
    double warmup_s;      /* building the resident index and bitmap */
    uint64_t ticks;               /* device batches launched */
    uint64_t max_tick;            /* most requests in one batch */
    uint64_t deferred_by_memory;  /* requests held back at least once by the memory budget */
    uint64_t peak_inflight_bytes; /* largest working set admitted at once */
} voss_amd_serve_stats_t;

/* Fill cfg with the defaults. */
int voss_amd_server_default_config(voss_amd_server_config_t* cfg);
/* Start (or restart) the server; NULL cfg = defaults.  Builds the resident state. */
int voss_amd_server_start(const voss_amd_server_config_t* cfg);
int voss_amd_server_stop(void);
/* Enqueue one request (arrival = now); *ticket identifies it for _wait. */
int voss_amd_server_submit(const voss_amd_request_t* req, uint64_t* ticket);
/* Block until the ticket's response is ready (each ticket is waited once).
 * A wait that has begun keeps its server alive until it returns, even if
 * another thread stops or replaces the server meanwhile (a server answers
 * every request it accepted before it goes away).  A ticket issued by an
 * earlier server returns VOSS_ERR_INVALID_ARG. */
int voss_amd_server_wait(uint64_t ticket, voss_amd_response_t* out);
/* Open-loop replay: request i is submitted arrival_s[i] seconds after the
 * start; returns every response and the latency/throughput summary. */
int voss_amd_server_replay(const voss_amd_request_t* reqs, const double* arrival_s, uint64_t n,
                           voss_amd_response_t* out, voss_amd_serve_stats_t* stats);

/* Engine configuration (the autotuner writes the same fields).  kind 0 is
 * used by sweeps (prime_count, Context, full statistics, bitmaps), kind 1 by
 * batched ranges; -1 sets both.  seg_bytes: bitmap bytes per LDS segment
 * (x30 integers; a multiple of 1 KiB x waves, at most 144 KiB); block: 512
 * or 1024 threads; wgs_per_cu: persistent workgroups per CU (LDS permitting);
 * tier1_p: primes below it are shared by a whole wave.  At start-up the
 * engine reads "<sweep|batch> seg block wgs tier1_p" lines from
 * $VOSS_TUNE_FILE or tuned.cfg next to libvoss.so. */
int voss_amd_set_config(uint32_t seg_bytes, int block, int wgs_per_cu, uint32_t tier1_p);
int voss_amd_get_config(uint32_t* seg_bytes, int* block, int* wgs_per_cu, uint32_t* tier1_p);
int voss_amd_set_config_for(int kind, uint32_t seg_bytes, int block, int wgs_per_cu, uint32_t tier1_p);
int voss_amd_get_config_for(int kind, uint32_t* seg_bytes, int* block, int* wgs_per_cu, uint32_t* tier1_p);
/* Device time of the last engine call (HIP events on the engine stream). */
int voss_amd_last_kernel_ms(double* sieve_ms, double* stitch_ms, double* total_ms, uint64_t* segments);
/* Measurement aid (no counterpart in the reference): the LDS rate this
 * device sustains for the sieve's access kinds, as lane-operations per clock
 * per CU at the shader clock measured during the run.  One 1024-thread
 * workgroup per CU over a 128 KiB LDS array, 8 * iters accesses per thread;
 * kind 0: ds_and_b32 at conflict-free addresses (consecutive dwords across
 * the lanes), 1: ds_and_b32 at pseudo-random addresses, 2: ds_read_b32
 * conflict-free, 3: ds_read_b32 random. */
int voss_amd_lds_probe(int kind, uint32_t iters, double* ops_per_clk_cu, double* clock_hz, double* ms);
/* Compute units and architecture name of the device in use. */
int voss_amd_device_info(int* num_cus, char* arch, int arch_len);
int voss_amd_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif /* VOSS_AMD_H */
