// kernels.h — launch entry points of the gfx950 kernels (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace voss {

struct StitchArgs {
  const SegSummary* summaries;
  const uint32_t* wave_gaps;  // per segment and wave: (dprev << 16) | largest gap
  uint32_t waves;             // waves per segment of the sweep
  uint32_t nseg;
  uint32_t seg_bytes;
  uint64_t k_base;
  uint64_t seed_last;   // last prime before the first segment (5)
  uint32_t seed_max;    // largest gap among the seeds (2)
  uint32_t large_threshold;
  unsigned long long* hist;
  GapRec* large;
  unsigned int* n_large;
  uint32_t large_cap;
  GapRec* rec;
  unsigned int* n_rec;
  uint32_t rec_cap;
  RescanItem* rescan;
  unsigned int* n_rescan;
  uint32_t rescan_cap;
  unsigned long long* final_out;  // [0] last prime, [1] largest gap
};

struct RescanArgs {
  const RescanItem* items;
  const unsigned int* n_items;  // device-side item count (capped at items_cap)
  uint32_t items_cap;
  GapRec* out;
  unsigned int* n_out;
  uint32_t out_cap;
};

struct BatchStitchArgs {
  const QueryDesc* queries;
  const SegSummary* summaries;
  uint32_t nq;
  const uint32_t* nq_dev;  // if set, the query count is read here
  uint32_t seg_bytes;
  unsigned int* qhist;
  unsigned long long* first;
  unsigned long long* last;
};

// Dynamic LDS bytes the segment kernels need for a segment of seg_bytes.
uint32_t lds_bytes_for(uint32_t seg_bytes);

// Raise the dynamic-LDS limit of every segment kernel to lds_bytes.
hipError_t prepare_kernels(uint32_t lds_bytes);

hipError_t launch_sieve(int mode, int block, const SieveArgs& A, int grid, hipStream_t st);
hipError_t launch_stitch(const StitchArgs& A, hipStream_t st);
// grid workgroups loop over the *R.n_items flagged wave ranges
hipError_t launch_rescan(int block, const SieveArgs& A, const RescanArgs& R, int grid, hipStream_t st);
hipError_t launch_batch_stitch(const BatchStitchArgs& A, uint32_t nq_cap, hipStream_t st);
// k-th prime (1-based ranks[i]) of sweep segment segs[i] for i < *n_dev; one
// workgroup per request, n_cap workgroups launched.
hipError_t launch_select(int block, const SieveArgs& A, const uint32_t* segs, const uint32_t* ranks,
                         unsigned long long* out, const uint32_t* n_dev, uint32_t n_cap, hipStream_t st);
// LDS rate probe (kind: bit 0 random addresses, bit 1 reads instead of ANDs):
// `blocks` 1024-thread workgroups, 8 * iters accesses per thread; clocks[0] /
// clocks[1] = shader-clock / 100 MHz real-time ticks of block 0.
hipError_t launch_lds_probe(int kind, uint32_t blocks, uint32_t iters, unsigned long long* clocks, uint32_t* sink,
                            hipStream_t st);
// Zero n 32-bit words (graph-safe fill, no memset node).
hipError_t launch_zero(uint32_t* p, uint64_t n, hipStream_t st);
// out[0] = add + sums[0].count + ... + sums[n-1].count (one workgroup).
hipError_t launch_sum_counts(const SegSummary* sums, uint32_t n, uint64_t add, unsigned long long* out,
                             hipStream_t st);

// Up to kMaxSpans word ranges zeroed (src null) or copied, in one launch: the
// sweep clears its counters and packs its results for one download this way
// instead of a memset or copy per buffer.
constexpr int kMaxSpans = 8;
struct CopySpan {
  const uint32_t* src;  // null: zero dst
  uint32_t* dst;
  uint64_t words;
};
struct SpanList {
  CopySpan s[kMaxSpans];
  uint32_t n;
};
hipError_t launch_spans(const SpanList& L, hipStream_t st);

// Bitmap -> ascending list of primes, in blocks of 4096 bitmap bytes:
// counts[b] = primes in block b; out[block_off[b] + i] = i-th prime of block b.
constexpr uint32_t kListBlockBytes = 4096;
hipError_t launch_list_count(const uint8_t* bm, uint64_t nbytes, uint32_t* counts, hipStream_t st);
hipError_t launch_list_write(const uint8_t* bm, uint64_t nbytes, uint64_t k_base,
                             const unsigned long long* block_off, uint64_t* out, hipStream_t st);

}  // namespace voss
