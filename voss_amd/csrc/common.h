// common.h — types shared by the host engine and the gfx950 kernels.
//
// Bitmap layout (the reference's, src/voss_w30_v7.cu:38-39): byte k of the
// global bitmap holds the eight wheel-30 candidates 30k + W30[o], o = 0..7, in
// bit o.  A segment is a run of consecutive bytes [k0, k0 + nbytes) sieved in
// one workgroup's LDS.  Values inside a segment are carried as u32 offsets
// from 30*k0 ("segment-relative values"); 0 means "no prime".
#pragma once
#include <stdint.h>

namespace voss {

// Primes 7..19 are removed by copying a periodic byte pattern (period
// 7*11*13*17*19 = 323,323 bytes) into LDS as the segment's initial bytes;
// 23..109 by AND-ing seven more patterns, three consecutive primes each
// (periods 23*29*31 = 20,677 bytes up to 103*107*109 = 1,201,289; the
// "read patterns") into every 16-byte chunk where the sieved segment is
// read: crossing-off never reads the bitmap, so the AND may come last.
// Crossing-off starts at 113, the sieving table's first prime.
constexpr int kPresieveGroups = 8;
constexpr uint32_t kPresievePeriods[kPresieveGroups] = {
    7u * 11u * 13u * 17u * 19u, 23u * 29u * 31u, 37u * 41u * 43u, 47u * 53u * 59u,
    61u * 67u * 71u,            73u * 79u * 83u, 89u * 97u * 101u, 103u * 107u * 109u};
constexpr uint32_t kFirstSievingPrime = 113;
// every presieved prime lies in the bitmap's bytes 0..3 (values below 120),
// which segment 0 keeps as their true bits
static_assert(kFirstSievingPrime <= 120, "presieved primes must lie below 120");
// Primes below this are sieved by tier 0 (a half-wave per prime and class).
constexpr uint32_t kTier0P = 160;
constexpr uint32_t kMaxSegBytes = 144u * 1024u;                  // LDS budget per segment
// LDS gap histogram: bins index gap/2 and cover every gap a batched query
// bins (2..512); longer gaps fall into the last bin and take an exact slow
// path.  Copies (one per 8-lane group) keep lanes apart on popular bins.
constexpr uint32_t kHistBins = 264;
constexpr uint32_t kHistCopies = 8;
constexpr uint32_t kGapBins = 50000;  // global histogram: index gap/2 for gap < 100000
constexpr uint32_t kQueryBins = 257;  // per-query histogram: [i] = gap 2i (i=1..256), [0] = other
constexpr uint32_t kPiTabStep = 1024; // sqrt table granularity

// One large gap (>= threshold).  Same field meaning as the reference's
// LargeGap (library/src/primes/gaps.cu:25-29): position = larger prime.
struct GapRec {
  uint64_t position;
  uint32_t gap;
  uint32_t seg;  // segment index (records) or 0
};

// Per-segment summary written by the sweep kernel.
struct SegSummary {
  uint32_t first;   // segment-relative value of the first prime, 0 if none
  uint32_t last;    // segment-relative value of the last prime, 0 if none
  uint32_t maxgap;  // largest gap between two primes of this segment, 0 if < 2 primes
  uint32_t count;   // primes in the segment
};

// A flagged wave range (the seg_bytes / waves bytes one wave of the sweep
// walked) whose records must be listed: every gap > threshold in order, each
// raising the threshold.  dprev = distance from the last prime before the
// range (in the same segment) to the range's first value, 0 if none.
struct RescanItem {
  uint32_t seg;
  uint32_t threshold;
  uint32_t wave;
  uint32_t dprev;
};

// Batched range query, flattened into segments by the host.
struct QueryDesc {
  uint64_t a, b;          // inclusive value range (a >= 7 here; 2, 3, 5 are host seeds)
  uint64_t k_lo, k_hi;    // byte range [k_lo, k_hi)
  uint32_t seg_first;     // index of the query's first segment in the work-list
  uint32_t nseg;
  uint32_t from_cache;    // 1: the segments' bytes come from the resident cache bitmap (no crossing-off)
  uint32_t pad;
};

enum KernelMode : int {
  kModeCount = 0,   // prime count only
  kModeStats = 1,   // sweep: count, tallies, histogram, summaries, large gaps
  kModeBatch = 2,   // batched ranges: per-query count, histogram, summaries
  kModeBitmap = 3,  // sweep: write the sieved, range-masked bytes to a global bitmap
};

// Arguments of the segment kernel (passed by value).
struct SieveArgs {
  // resident tables
  const uint32_t* primes;       // sieving primes >= 23, bits 24..26 = index of p mod 30
  const uint32_t* pi_tab;       // pi_tab[i] = #sieving primes <= i*kPiTabStep
  uint32_t pi_tab_len;
  const uint32_t* presieve;     // pattern bytes as u32 words, kPresieveGroups patterns
  uint32_t presieve_off[kPresieveGroups];  // word offset of each pattern
  uint32_t presieve_copy_words[kPresieveGroups];  // each pattern is stored 4 times, copy c shifted by c bytes, this many words apart
  uint32_t first_word;          // true prime bitmap of bytes 0..3 (values < 120)
  // geometry
  uint64_t k_base;              // sweep: byte index of segment 0
  uint64_t k_end;               // sweep: exclusive end byte
  uint64_t lo_val, hi_val;      // sweep: inclusive value bounds counted
  uint32_t seg_bytes;           // S (multiple of 16 * waves * 64... see host)
  uint32_t nseg;                // segments in this launch
  uint32_t seg_base;            // global index of this launch's segment 0
  uint32_t tier0_end;           // primes [0, tier0_end): a half-wave per (prime, class), lanes split the crossings
  uint32_t tier1_end;           // primes [tier0_end, tier1_end): one prime per wave, lanes split the segment
  uint32_t tier2_end;           // primes [tier1_end, tier2_end): one per lane, unrolled walk
  uint32_t tier3_end;           // primes [tier2_end, tier3_end): one per lane, <= 4 hits/class; rest (p >= seg_bytes): <= 1
  unsigned int* work;           // segment work counter of the batch, count and stats modes (zeroed before launch)
  const uint32_t* nseg_dev;     // batch mode: if set, the segment count is read here
  // outputs (mode dependent; unused ones may be null)
  unsigned long long* totals;   // [0] count, [1] pi_4_1 part, [2] pi_3_1 part
  SegSummary* summaries;        // indexed by global segment
  uint32_t* wave_gaps;          // stats mode: per global segment and wave, (dprev << 16) | largest gap
  unsigned long long* hist;     // kGapBins u64
  unsigned long long* buckets;  // per checkpoint bucket: count, c41, c31
  uint64_t cp_interval;         // 0 = no buckets
  uint32_t nbuckets;
  GapRec* large;
  unsigned int* n_large;
  uint32_t large_cap;
  uint32_t large_threshold;
  // batch mode
  const QueryDesc* queries;
  const uint32_t* seg_query;    // query index per segment
  unsigned int* qhist;          // kQueryBins u32 per query
  unsigned long long* qcount;   // per query
  // bitmap mode
  uint8_t* bitmap;              // byte (k - k_base) of the sieved range
  // batch mode: resident sieved bitmap of [7, cache_hi] from byte 0, read by
  // queries flagged from_cache instead of sieving their segments
  const uint8_t* cache;
};

}  // namespace voss
