/*
 * voss_oracle.c — CPU restatement of the voss wheel-30 sieve path.
 *
 * TEST INFRASTRUCTURE ONLY (see voss_oracle.h): the checker for the HIP
 * product path and the "port" CPU baseline.  Never linked into libvoss.so.
 *
 * Each function names the reference code it restates.  The sieve is a plain
 * multithreaded segmented sieve over the reference's byte-per-30 bitmap
 * (src/voss_w30_v7.cu:38-39; library/src/primes/sieve.cu:10-37 crossing rule:
 * multiples p*m with m >= p, one bit per (k, residue)), with per-thread
 * crossing offsets carried from segment to segment.  Statistics follow
 * library/src/primes/context.cu:212-559 (seeds, fix-ups, boundary stitching)
 * and gaps.cu:31-75 / chebyshev.cu:11-25 (per-gap and per-prime rules).
 */
#define _GNU_SOURCE
#include "voss_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static const uint32_t W30[8] = {1, 7, 11, 13, 17, 19, 23, 29};
static int RES2BIT[30];
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

static void init_tables(void) {
    for (int r = 0; r < 30; ++r) RES2BIT[r] = -1;
    for (int o = 0; o < 8; ++o) RES2BIT[W30[o]] = o;
}

static int resolve_threads(int threads) {
    if (threads > 0) return threads;
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

static uint64_t isqrt_u64(uint64_t n) {
    uint64_t r = (uint64_t)sqrt((double)n);
    while (r * r > n) --r;
    while ((r + 1) * (r + 1) <= n) ++r;
    return r;
}

/* ---- growable array of gaps ---- */
typedef struct { orc_gap_t* v; uint64_t n, cap; } gapvec;
static void gv_push(gapvec* g, uint64_t pos, uint32_t gap) {
    if (g->n == g->cap) {
        g->cap = g->cap ? g->cap * 2 : 16;
        g->v = (orc_gap_t*)realloc(g->v, g->cap * sizeof(orc_gap_t));
    }
    g->v[g->n].position = pos; g->v[g->n].gap = gap; g->v[g->n].pad = 0;
    g->n++;
}

/* ---- base primes (context.cu:107-121 restated: primes >= 7 up to limit) ---- */
static uint32_t* base_primes(uint64_t limit, size_t* count) {
    uint8_t* s = (uint8_t*)malloc(limit + 1);
    memset(s, 1, limit + 1);
    s[0] = 0; if (limit >= 1) s[1] = 0;
    for (uint64_t i = 2; i * i <= limit; ++i)
        if (s[i]) for (uint64_t j = i * i; j <= limit; j += i) s[j] = 0;
    size_t c = 0;
    for (uint64_t i = 7; i <= limit; ++i) c += s[i];
    uint32_t* out = (uint32_t*)malloc((c ? c : 1) * sizeof(uint32_t));
    size_t k = 0;
    for (uint64_t i = 7; i <= limit; ++i) if (s[i]) out[k++] = (uint32_t)i;
    free(s);
    *count = c;
    return out;
}

/* ---- crossing state: next byte index per (prime, class) ---- */
typedef struct {
    const uint32_t* p;
    size_t np;
    uint64_t* next;
    uint8_t* mask; /* byte mask with the class's bit cleared */
} cross_t;

/* First crossing of each class at or after byte K, never below p*p
 * (sieve rule of context.cu:354-388: n >= max(p^2, segment low)). */
static void cross_init(cross_t* cs, const uint32_t* p, size_t np, uint64_t K) {
    cs->p = p; cs->np = np;
    cs->next = (uint64_t*)malloc((np ? np : 1) * 8 * sizeof(uint64_t));
    cs->mask = (uint8_t*)malloc((np ? np : 1) * 8);
    for (size_t i = 0; i < np; ++i) {
        uint64_t P = p[i];
        uint64_t ms = (30 * K + P - 1) / P;
        if (ms < P) ms = P;
        for (int c = 0; c < 8; ++c) {
            uint64_t m = ms + (W30[c] + 30 - ms % 30) % 30;
            uint64_t v = P * m;
            cs->next[i * 8 + c] = v / 30;
            cs->mask[i * 8 + c] = (uint8_t)~(1u << RES2BIT[v % 30]);
        }
    }
}
static void cross_free(cross_t* cs) { free(cs->next); free(cs->mask); }

/* Sieve bytes [s_lo, s_hi) into seg (all candidates start set; 1 is cleared). */
static void sieve_segment(cross_t* cs, uint8_t* seg, uint64_t s_lo, uint64_t s_hi) {
    uint64_t len = s_hi - s_lo;
    memset(seg, 0xFF, len);
    if (s_lo == 0) seg[0] &= 0xFE; /* the number 1 (context.cu:404-410) */
    for (size_t i = 0; i < cs->np; ++i) {
        uint64_t P = cs->p[i];
        if ((P * P) / 30 >= s_hi) break;
        uint64_t* nx = cs->next + i * 8;
        const uint8_t* mk = cs->mask + i * 8;
        for (int c = 0; c < 8; ++c) {
            uint64_t j = nx[c];
            uint8_t m = mk[c];
            while (j < s_hi) { seg[j - s_lo] &= m; j += P; }
            nx[c] = j;
        }
    }
}

/* ---- statistics accumulation ---- */
typedef struct {
    /* order-independent (per thread) */
    uint64_t* hist;       /* ORC_MAX_GAP, may be NULL */
    uint64_t hist_over;
    uint64_t* bucket;     /* 5 per checkpoint bucket, may be NULL */
    uint64_t nbuckets, interval;
} thread_acc;

typedef struct {
    uint64_t count, first, last, maxgap;
    gapvec large, recs;
} chunk_res;

typedef struct {
    /* job */
    uint64_t k_begin, k_end;     /* byte range */
    uint64_t lo_val, hi_val;     /* inclusive value range counted */
    uint64_t chunk_bytes, nchunks;
    const uint32_t* primes; size_t np;
    int stats;                   /* 0 = count only */
    uint32_t large_threshold;
    uint64_t interval, nbuckets;
    /* outputs */
    chunk_res* chunks;
    atomic_ulong next_chunk;
    thread_acc* accs;
} job_t;

typedef struct { job_t* job; int tid; } worker_arg;

#define SEG_BYTES (64 * 1024)

static inline void tally(thread_acc* ta, uint64_t v) {
    if (!ta->bucket) return;
    uint64_t b = (v - 1) / ta->interval;
    if (b >= ta->nbuckets) b = ta->nbuckets - 1;
    uint64_t* t = ta->bucket + 5 * b;
    t[0]++;
    if ((v & 3) == 1) t[1]++; else t[2]++;
    uint64_t r3 = v % 3;
    if (r3 == 1) t[3]++; else if (r3 == 2) t[4]++;
}

static void* worker(void* argp) {
    worker_arg* wa = (worker_arg*)argp;
    job_t* J = wa->job;
    thread_acc* ta = &J->accs[wa->tid];
    uint8_t* seg = (uint8_t*)malloc(SEG_BYTES);
    for (;;) {
        uint64_t c = atomic_fetch_add(&J->next_chunk, 1);
        if (c >= J->nchunks) break;
        chunk_res* R = &J->chunks[c];
        memset(R, 0, sizeof(*R));
        uint64_t cb = J->k_begin + c * J->chunk_bytes;
        uint64_t ce = cb + J->chunk_bytes;
        if (ce > J->k_end) ce = J->k_end;
        /* primes needed: p*p <= 30*ce */
        size_t np = 0;
        while (np < J->np && (uint64_t)J->primes[np] * J->primes[np] < 30 * ce) ++np;
        cross_t cs;
        cross_init(&cs, J->primes, np, cb);
        uint64_t prev = 0, lmax = 0;
        for (uint64_t s = cb; s < ce; s += SEG_BYTES) {
            uint64_t se = s + SEG_BYTES; if (se > ce) se = ce;
            sieve_segment(&cs, seg, s, se);
            for (uint64_t j = s; j < se; ++j) {
                uint32_t b = seg[j - s];
                while (b) {
                    int o = __builtin_ctz(b);
                    b &= b - 1;
                    uint64_t v = 30 * j + W30[o];
                    if (v < J->lo_val) continue;
                    if (v > J->hi_val) { b = 0; break; }
                    R->count++;
                    if (!J->stats) continue;
                    if (!R->first) R->first = v;
                    tally(ta, v);
                    if (prev) {
                        uint64_t g = v - prev;
                        if (ta->hist) {
                            if (g < ORC_MAX_GAP) ta->hist[g]++; else ta->hist_over++;
                        }
                        if (g >= J->large_threshold) gv_push(&R->large, v, (uint32_t)g);
                        if (g > lmax) { gv_push(&R->recs, v, (uint32_t)g); lmax = g; }
                    }
                    prev = v;
                }
            }
        }
        R->last = prev;
        R->maxgap = lmax;
        cross_free(&cs);
    }
    free(seg);
    return NULL;
}

/* Run the chunked sieve over bytes [k_begin, k_end) counting values in [lo, hi]. */
static job_t* run_job(uint64_t lo, uint64_t hi, int stats, uint32_t thr_large,
                      uint64_t interval, uint64_t nbuckets, int threads, uint64_t hist_len) {
    pthread_once(&g_once, init_tables);
    job_t* J = (job_t*)calloc(1, sizeof(job_t));
    J->k_begin = lo / 30;
    J->k_end = hi / 30 + 1;
    J->lo_val = lo; J->hi_val = hi;
    J->stats = stats;
    J->large_threshold = thr_large;
    J->interval = interval; J->nbuckets = nbuckets;
    uint64_t nb = J->k_end - J->k_begin;
    int T = resolve_threads(threads);
    uint64_t cb = nb / ((uint64_t)T * 8) + 1;
    if (cb < SEG_BYTES) cb = SEG_BYTES;
    if (cb > 64ull * SEG_BYTES) cb = 64ull * SEG_BYTES;
    J->chunk_bytes = cb;
    J->nchunks = (nb + cb - 1) / cb;
    J->primes = base_primes(isqrt_u64(hi) + 1, &J->np);
    J->chunks = (chunk_res*)calloc(J->nchunks ? J->nchunks : 1, sizeof(chunk_res));
    atomic_init(&J->next_chunk, 0);
    J->accs = (thread_acc*)calloc(T, sizeof(thread_acc));
    for (int t = 0; t < T; ++t) {
        if (stats && hist_len) J->accs[t].hist = (uint64_t*)calloc(hist_len, sizeof(uint64_t));
        if (stats && nbuckets) J->accs[t].bucket = (uint64_t*)calloc(5 * nbuckets, sizeof(uint64_t));
        J->accs[t].nbuckets = nbuckets; J->accs[t].interval = interval;
    }
    pthread_t* th = (pthread_t*)malloc(T * sizeof(pthread_t));
    worker_arg* wa = (worker_arg*)malloc(T * sizeof(worker_arg));
    for (int t = 0; t < T; ++t) { wa[t].job = J; wa[t].tid = t; pthread_create(&th[t], NULL, worker, &wa[t]); }
    for (int t = 0; t < T; ++t) pthread_join(th[t], NULL);
    free(th); free(wa);
    J->nchunks = J->nchunks; /* all done */
    return J;
}

static void free_job(job_t* J, int T) {
    for (uint64_t c = 0; c < J->nchunks; ++c) { free(J->chunks[c].large.v); free(J->chunks[c].recs.v); }
    for (int t = 0; t < T; ++t) { free(J->accs[t].hist); free(J->accs[t].bucket); }
    free(J->accs); free(J->chunks); free((void*)J->primes); free(J);
}

static int cmp_large(const void* a, const void* b) {
    const orc_gap_t* x = (const orc_gap_t*)a; const orc_gap_t* y = (const orc_gap_t*)b;
    if (x->gap != y->gap) return x->gap > y->gap ? -1 : 1;
    return x->position < y->position ? -1 : (x->position > y->position);
}

/* Full statistics: restates compute_full_impl (context.cu:212-559) with
 * one ordered merge of per-chunk first/last/records in place of the
 * per-segment host loop. */
int orc_full_stats(uint64_t N, uint64_t cp_interval, uint32_t large_threshold,
                   int threads, orc_stats_t* out, uint64_t* hist,
                   uint64_t* cp_out, orc_gap_t* large, uint64_t large_cap,
                   orc_gap_t* records, uint64_t rec_cap) {
    if (!out || N < 2) return 1;
    memset(out, 0, sizeof(*out));
    if (hist) memset(hist, 0, ORC_MAX_GAP * sizeof(uint64_t));
    uint64_t ncp = cp_interval ? N / cp_interval : 0;
    if (cp_interval && cp_interval < 30) return 1;

    /* small primes handled as seeds */
    static const uint64_t seeds[3] = {2, 3, 5};
    gapvec recs = {0}, lg = {0};
    uint64_t nbuckets = cp_interval ? (N + cp_interval - 1) / cp_interval : 0;
    uint64_t* bucket = nbuckets ? (uint64_t*)calloc(5 * nbuckets, sizeof(uint64_t)) : NULL;
    uint64_t prev = 0, rmax = 0, count = 0;
    for (int i = 0; i < 3 && seeds[i] <= N; ++i) {
        uint64_t v = seeds[i];
        count++;
        if (bucket) {
            uint64_t* t = bucket + 5 * ((v - 1) / cp_interval);
            t[0]++;
            if (v != 2) { if ((v & 3) == 1) t[1]++; else t[2]++; }
            if (v % 3 == 1) t[3]++; else if (v % 3 == 2) t[4]++;
        }
        if (prev) {
            uint64_t g = v - prev;
            if (hist) hist[g]++;
            if (g > rmax) { gv_push(&recs, v, (uint32_t)g); rmax = g; }
        }
        prev = v;
    }
    /* reference mod-4 convention: +1 for 5 (class 1), +1 for 3 (class 3),
     * applied only when N >= 7 (context.cu:228-247, 520-524) */
    uint64_t c41 = 0, c43 = 0;
    if (N >= 7) { c41 = 1; c43 = 1; }
    uint64_t c31 = 0, c32 = (N >= 5) ? 2 : (N >= 2 ? 1 : 0);

    int T = resolve_threads(threads);
    if (N >= 7) {
        job_t* J = run_job(7, N, 1, large_threshold, cp_interval ? cp_interval : N,
                           nbuckets ? nbuckets : 1, threads, ORC_MAX_GAP);
        for (int t = 0; t < T; ++t) {
            if (hist) for (uint64_t g = 0; g < ORC_MAX_GAP; ++g) hist[g] += J->accs[t].hist[g];
            out->hist_overflow += J->accs[t].hist_over;
            if (bucket) for (uint64_t b = 0; b < 5 * nbuckets; ++b) bucket[b] += J->accs[t].bucket[b];
            uint64_t* tb = J->accs[t].bucket;
            for (uint64_t b = 0; b < (nbuckets ? nbuckets : 1); ++b) {
                c41 += tb[5 * b + 1]; c43 += tb[5 * b + 2];
                c31 += tb[5 * b + 3]; c32 += tb[5 * b + 4];
            }
        }
        for (uint64_t c = 0; c < J->nchunks; ++c) {
            chunk_res* R = &J->chunks[c];
            if (!R->count) continue;
            count += R->count;
            /* boundary gap: histogram, large list and records
             * (context.cu:472-478 stitches the histogram only; the large list
             * here is the true list — see orc docs) */
            uint64_t g = R->first - prev;
            if (hist) { if (g < ORC_MAX_GAP) hist[g]++; else out->hist_overflow++; }
            if (g >= large_threshold) gv_push(&lg, R->first, (uint32_t)g);
            if (g > rmax) { gv_push(&recs, R->first, (uint32_t)g); rmax = g; }
            for (uint64_t i = 0; i < R->recs.n; ++i)
                if (R->recs.v[i].gap > rmax) { gv_push(&recs, R->recs.v[i].position, R->recs.v[i].gap); rmax = R->recs.v[i].gap; }
            for (uint64_t i = 0; i < R->large.n; ++i) gv_push(&lg, R->large.v[i].position, R->large.v[i].gap);
            prev = R->last;
        }
        free_job(J, T);
    }
    out->prime_count = count;
    out->pi_4_1 = c41; out->pi_4_3 = c43;
    out->pi_3_1 = c31; out->pi_3_2 = c32;
    out->last_prime = prev;
    out->max_gap = rmax;
    out->n_large = lg.n;
    out->n_records = recs.n;
    if (lg.n > 1) qsort(lg.v, lg.n, sizeof(orc_gap_t), cmp_large);  /* lg.v is NULL while empty */
    if (large) for (uint64_t i = 0; i < lg.n && i < large_cap; ++i) large[i] = lg.v[i];
    if (records) for (uint64_t i = 0; i < recs.n && i < rec_cap; ++i) records[i] = recs.v[i];
    if (cp_out && bucket) {
        uint64_t acc[5] = {0, 0, 0, 0, 0};
        for (uint64_t j = 0; j < ncp; ++j) {
            for (int f = 0; f < 5; ++f) acc[f] += bucket[5 * j + f];
            for (int f = 0; f < 5; ++f) cp_out[5 * j + f] = acc[f];
        }
    }
    free(bucket); free(lg.v); free(recs.v);
    return 0;
}

uint64_t orc_prime_count(uint64_t N, int threads) {
    if (N < 2) return 0;
    uint64_t count = (N >= 5) ? 3 : (N >= 3 ? 2 : 1);
    if (N < 7) return count;
    int T = resolve_threads(threads);
    job_t* J = run_job(7, N, 0, 0xFFFFFFFFu, N, 1, threads, 0);
    for (uint64_t c = 0; c < J->nchunks; ++c) count += J->chunks[c].count;
    free_job(J, T);
    return count;
}

/* Range statistics over true primes in [a, b]; no reference counterpart
 * (new batched-range output), stated in voss_oracle.h. */
int orc_range_stats(uint64_t a, uint64_t b, int threads, uint64_t* count,
                    uint64_t* first, uint64_t* last, uint64_t* hist257) {
    if (!count || !first || !last || !hist257) return 1;
    *count = 0; *first = 0; *last = 0;
    memset(hist257, 0, 257 * sizeof(uint64_t));
    if (b < a || b < 2) return 0;
    uint64_t prev = 0;
    #define RANGE_ADD(v) do { uint64_t vv_ = (v); (*count)++; if (!*first) *first = vv_; \
        if (prev) { uint64_t g_ = vv_ - prev; if (g_ >= 2 && g_ <= 512 && !(g_ & 1)) hist257[g_ / 2]++; else hist257[0]++; } \
        prev = vv_; } while (0)
    static const uint64_t seeds[3] = {2, 3, 5};
    for (int i = 0; i < 3; ++i) if (seeds[i] >= a && seeds[i] <= b) RANGE_ADD(seeds[i]);
    uint64_t lo = a < 7 ? 7 : a;
    if (b >= lo) {
        int T = resolve_threads(threads);
        job_t* J = run_job(lo, b, 1, 0xFFFFFFFFu, 0, 0, threads, ORC_MAX_GAP);
        for (int t = 0; t < T; ++t)
            for (uint64_t g = 1; g < ORC_MAX_GAP; ++g) {
                uint64_t h = J->accs[t].hist[g];
                if (!h) continue;
                if (g >= 2 && g <= 512 && !(g & 1)) hist257[g / 2] += h; else hist257[0] += h;
            }
        for (int t = 0; t < T; ++t) hist257[0] += J->accs[t].hist_over;
        for (uint64_t c = 0; c < J->nchunks; ++c) {
            chunk_res* R = &J->chunks[c];
            if (!R->count) continue;
            /* chunk-internal gaps are already in the thread histograms */
            if (prev) {
                uint64_t g = R->first - prev;
                if (g >= 2 && g <= 512 && !(g & 1)) hist257[g / 2]++; else hist257[0]++;
            }
            if (!*first) *first = R->first;
            *count += R->count;
            prev = R->last;
        }
        free_job(J, T);
    }
    *last = prev;
    #undef RANGE_ADD
    return 0;
}

uint64_t orc_primes_in_range(uint64_t a, uint64_t b, uint64_t* out, uint64_t cap) {
    pthread_once(&g_once, init_tables);
    uint64_t n = 0;
    if (b < a || b < 2) return 0;
    static const uint64_t seeds[3] = {2, 3, 5};
    for (int i = 0; i < 3; ++i)
        if (seeds[i] >= a && seeds[i] <= b) { if (n < cap) out[n] = seeds[i]; n++; }
    uint64_t lo = a < 7 ? 7 : a;
    if (b < lo) return n;
    size_t np; uint32_t* pr = base_primes(isqrt_u64(b) + 1, &np);
    uint64_t kb = lo / 30, ke = b / 30 + 1;
    cross_t cs; cross_init(&cs, pr, np, kb);
    uint8_t* seg = (uint8_t*)malloc(SEG_BYTES);
    for (uint64_t s = kb; s < ke; s += SEG_BYTES) {
        uint64_t se = s + SEG_BYTES; if (se > ke) se = ke;
        sieve_segment(&cs, seg, s, se);
        for (uint64_t j = s; j < se; ++j) {
            uint32_t m = seg[j - s];
            while (m) {
                int o = __builtin_ctz(m); m &= m - 1;
                uint64_t v = 30 * j + W30[o];
                if (v < lo || v > b) continue;
                if (n < cap) out[n] = v;
                n++;
            }
        }
    }
    free(seg); cross_free(&cs); free(pr);
    return n;
}

/* ---- Miller-Rabin: restates library/src/core/miller_rabin.cpp:26-51 ---- */
static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) {
    return (uint64_t)((unsigned __int128)a * b % m);
}
static uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
    uint64_t r = 1;
    a %= m;
    while (e) { if (e & 1) r = mulmod(r, a, m); a = mulmod(a, a, m); e >>= 1; }
    return r;
}
int orc_is_prime_mr(uint64_t n) {
    if (n < 2) return 0;
    if (n == 2 || n == 3) return 1;
    if ((n & 1) == 0) return 0;
    uint64_t d = n - 1; int r = 0;
    while ((d & 1) == 0) { d >>= 1; r++; }
    static const uint64_t wit[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (int i = 0; i < 12; ++i) {
        uint64_t a = wit[i];
        if (a >= n) continue;
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int k = 0; k < r - 1; ++k) {
            x = mulmod(x, x, n);
            if (x == n - 1) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}

/* ---- nth prime: the value nth.cu:40-103 returns, found by counting ---- */
int orc_nth_prime(uint64_t n, uint64_t* out) {
    if (!out || n == 0) return 1;
    static const uint64_t small[6] = {2, 3, 5, 7, 11, 13};
    if (n <= 6) { *out = small[n - 1]; return 0; }
    double nf = (double)n;
    /* Rosser: p_n < n(ln n + ln ln n) for n >= 6 */
    uint64_t hi = (uint64_t)(nf * (log(nf) + log(log(nf)))) + 64;
    pthread_once(&g_once, init_tables);
    int T = resolve_threads(0);
    job_t* J = run_job(7, hi, 0, 0xFFFFFFFFu, hi, 1, T, 0);
    uint64_t cum = 3;
    uint64_t c;
    for (c = 0; c < J->nchunks; ++c) {
        if (cum + J->chunks[c].count >= n) break;
        cum += J->chunks[c].count;
    }
    if (c == J->nchunks) { free_job(J, T); return 2; }
    uint64_t kb = J->k_begin + c * J->chunk_bytes;
    uint64_t ke = kb + J->chunk_bytes; if (ke > J->k_end) ke = J->k_end;
    free_job(J, T);
    uint64_t lo = 30 * kb; if (lo < 7) lo = 7;
    uint64_t hv = 30 * ke - 1; if (hv > hi) hv = hi;
    uint64_t need = n - cum;
    uint64_t* buf = (uint64_t*)malloc(need * sizeof(uint64_t));
    uint64_t got = orc_primes_in_range(lo, hv, buf, need);
    int rc = 2;
    if (got >= need) { *out = buf[need - 1]; rc = 0; }
    free(buf);
    return rc;
}

/* ---- next/prev: restate next_prev.cu:22-83 (MR scan, 1e8 step cap) ---- */
int orc_next_prime(uint64_t x, uint64_t* out) {
    if (!out) return 6;
    if (x < 2) { *out = 2; return 0; }
    uint64_t c = x + 1;
    if (c == 2) { *out = 2; return 0; }
    if ((c & 1) == 0) c++;
    for (uint64_t steps = 0; steps < 100000000ull; steps += 2, c += 2)
        if (orc_is_prime_mr(c)) { *out = c; return 0; }
    return 2;
}
int orc_prev_prime(uint64_t x, uint64_t* out) {
    if (!out) return 6;
    if (x <= 2) return 1;
    if (x == 3) { *out = 2; return 0; }
    uint64_t c = x - 1;
    if (c == 2) { *out = 2; return 0; }
    if ((c & 1) == 0) c--;
    for (uint64_t steps = 0; steps < 100000000ull && c >= 3; steps += 2, c -= 2)
        if (orc_is_prime_mr(c)) { *out = c; return 0; }
    if (c < 2) { *out = 2; return 0; }
    return 2;
}

/* ---- Goldbach ---- */
/* goldbach.cpp:21-43: n == 4 -> (2,2); else odd p1 from 3 to n/2 with MR. */
uint64_t orc_goldbach_count_mr(uint64_t n) {
    if (n < 4 || (n & 1)) return 0;
    if (n == 4) return 1;
    uint64_t c = 0;
    for (uint64_t p = 3; p <= n / 2; p += 2)
        if (orc_is_prime_mr(p) && orc_is_prime_mr(n - p)) c++;
    return c;
}

/* odd-number prime bitmap: bit i <-> 2i+1 (2 handled by callers) */
typedef struct { uint64_t* w; uint64_t limit; } oddbm;
static oddbm oddbm_build(uint64_t limit) {
    oddbm B; B.limit = limit;
    uint64_t nbits = limit / 2 + 1;
    B.w = (uint64_t*)calloc(nbits / 64 + 2, sizeof(uint64_t));
    uint64_t cap = limit / 2 + 16;
    uint64_t* pr = (uint64_t*)malloc(cap * sizeof(uint64_t));
    uint64_t n = orc_primes_in_range(3, limit, pr, cap);
    for (uint64_t i = 0; i < n && i < cap; ++i) { uint64_t b = pr[i] >> 1; B.w[b >> 6] |= 1ull << (b & 63); }
    free(pr);
    return B;
}
static inline int oddbm_get(const oddbm* B, uint64_t v) {
    uint64_t b = v >> 1; return (int)((B->w[b >> 6] >> (b & 63)) & 1);
}

uint64_t orc_goldbach_count_sieve(uint64_t n) {
    if (n < 4 || (n & 1)) return 0;
    if (n == 4) return 1;
    oddbm B = oddbm_build(n);
    uint64_t c = 0;
    for (uint64_t p = 3; p <= n / 2; p += 2) if (oddbm_get(&B, p) && oddbm_get(&B, n - p)) c++;
    free(B.w);
    return c;
}

typedef struct { const uint64_t* ns; uint64_t k; uint64_t* out; const oddbm* B; atomic_ulong next; } gb_job;
static void* gb_worker(void* a) {
    gb_job* J = (gb_job*)a;
    for (;;) {
        uint64_t i = atomic_fetch_add(&J->next, 1);
        if (i >= J->k) break;
        uint64_t n = J->ns[i], c = 0;
        if (n == 4) c = 1;
        else if (n >= 6 && !(n & 1))
            for (uint64_t p = 3; p <= n / 2; p += 2) if (oddbm_get(J->B, p) && oddbm_get(J->B, n - p)) c++;
        J->out[i] = c;
    }
    return NULL;
}
int orc_goldbach_counts(const uint64_t* ns, uint64_t k, uint64_t* out, int threads) {
    uint64_t nmax = 4;
    for (uint64_t i = 0; i < k; ++i) if (ns[i] > nmax) nmax = ns[i];
    oddbm B = oddbm_build(nmax);
    gb_job J; J.ns = ns; J.k = k; J.out = out; J.B = &B; atomic_init(&J.next, 0);
    int T = resolve_threads(threads);
    pthread_t* th = (pthread_t*)malloc(T * sizeof(pthread_t));
    for (int t = 0; t < T; ++t) pthread_create(&th[t], NULL, gb_worker, &J);
    for (int t = 0; t < T; ++t) pthread_join(th[t], NULL);
    free(th); free(B.w);
    return 0;
}

/* goldbach_upto.cu:25-46: for each even n in [4, N], does some prime
 * p <= n/2 have n - p prime?  Counterexamples ascending. */
#define GBT_BLOCK (1ull << 20)
typedef struct {
    const oddbm* B; uint64_t N, pmax, keep;
    uint64_t** found; uint64_t* nfound; atomic_ulong next; uint64_t blocks;
} gbt_job;
static void* gbt_worker(void* a) {
    gbt_job* J = (gbt_job*)a;
    for (;;) {
        uint64_t blk = atomic_fetch_add(&J->next, 1);
        if (blk >= J->blocks) break;
        uint64_t n0 = 6 + blk * GBT_BLOCK * 2;
        uint64_t nf = 0, cap = 0;
        uint64_t* list = NULL;
        for (uint64_t i = 0; i < GBT_BLOCK; ++i) {
            uint64_t n = n0 + 2 * i;
            if (n > J->N) break;
            int ok = 0;
            for (uint64_t p = 3; p <= n / 2 && p <= J->pmax; p += 2)
                if (oddbm_get(J->B, p) && oddbm_get(J->B, n - p)) { ok = 1; break; }
            if (ok) continue;
            if (nf < J->keep) {  /* the block's first `keep` counterexamples, ascending */
                if (nf == cap) {
                    cap = cap ? 2 * cap : 16;
                    list = (uint64_t*)realloc(list, cap * sizeof(uint64_t));
                }
                list[nf] = n;
            }
            nf++;
        }
        J->found[blk] = list;
        J->nfound[blk] = nf;
    }
    return NULL;
}
uint64_t orc_goldbach_test_upto(uint64_t N, uint64_t* ce, uint64_t max_ce, int threads) {
    return orc_goldbach_test_upto_pmax(N, UINT64_MAX, ce, max_ce, threads);
}

/* Same scan with partners p restricted to p <= pmax: a planted defect whose
 * counterexamples a brute force can list, so the scan and its reporting are
 * pinned even though the true answer below 1e9 is "none". */
uint64_t orc_goldbach_test_upto_pmax(uint64_t N, uint64_t pmax, uint64_t* ce, uint64_t max_ce, int threads) {
    if (N < 6) return 0; /* 4 = 2 + 2 */
    oddbm B = oddbm_build(N);
    gbt_job J; J.B = &B; J.N = N; J.pmax = pmax; J.keep = max_ce;
    J.blocks = ((N - 6) / 2) / GBT_BLOCK + 1;
    J.found = (uint64_t**)calloc(J.blocks, sizeof(uint64_t*));
    J.nfound = (uint64_t*)calloc(J.blocks, sizeof(uint64_t));
    atomic_init(&J.next, 0);
    int T = resolve_threads(threads);
    pthread_t* th = (pthread_t*)malloc(T * sizeof(pthread_t));
    for (int t = 0; t < T; ++t) pthread_create(&th[t], NULL, gbt_worker, &J);
    for (int t = 0; t < T; ++t) pthread_join(th[t], NULL);
    uint64_t n = 0;
    for (uint64_t b = 0; b < J.blocks; ++b) {
        for (uint64_t i = 0; i < J.nfound[b]; ++i) {
            if (i < J.keep && n < max_ce && ce) ce[n] = J.found[b][i];
            n++;
        }
        free(J.found[b]);
    }
    free(th); free(J.found); free(J.nfound); free(B.w);
    return n;
}

/* sophie_germain_upto.cu:22-33 with context.cu:765-888: count p <= N with
 * p and 2p+1 prime. */
uint64_t orc_sophie_germain(uint64_t N) {
    if (N < 2) return 0;
    oddbm B = oddbm_build(2 * N + 1);
    uint64_t c = (N >= 2) ? 1 : 0; /* 2 -> 5 */
    for (uint64_t p = 3; p <= N; p += 2) if (oddbm_get(&B, p) && oddbm_get(&B, 2 * p + 1)) c++;
    free(B.w);
    return c;
}
