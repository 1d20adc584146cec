// engine.h — host-side orchestration of the gfx950 sieve (one process, one GPU).
//
// The Engine owns the device state that stays resident between calls: the
// sieving-prime table, the sqrt lookup table, the presieve pattern, one
// stream and reusable scratch buffers.  Every public method throws
// voss::Error on failure; the C ABI layer maps errors to VOSS_ERR_* codes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.h"

namespace voss {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Launch configuration chosen by the autotuner (or set by the caller).
struct TuneConfig {
  uint32_t seg_bytes = 144 * 1024;  // bytes of bitmap per segment (x30 integers)
  int block = 1024;                 // threads per workgroup
  int wgs_per_cu = 1;               // persistent workgroups per CU
  uint32_t tier1_p = 1536;          // primes below this: one prime per half-wave
};

struct GapEntry {
  uint64_t position;  // larger prime
  uint32_t gap;
};

// Result of a full-statistics sweep over [2, N] (N >= 7).
struct SweepStats {
  uint64_t prime_count = 0;
  uint64_t pi_4_1 = 0, pi_4_3 = 0, pi_3_1 = 0, pi_3_2 = 0;
  uint64_t last_prime = 0;
  uint64_t max_gap = 0;
  std::vector<uint64_t> hist;                // index = gap, up to the largest gap (all later counts are 0)
  std::vector<uint64_t> checkpoints;         // 5 per checkpoint: pi, c41, c43, c31, c32
  std::vector<GapEntry> large;               // gaps >= threshold, (gap desc, position asc)
  std::vector<GapEntry> records;             // maximal gaps, ascending position
  uint64_t large_total = 0;                  // found (may exceed the stored list)
};

// Laid out as the C ABI's voss_amd_range_t (pad included), so batch results
// can be written straight into a caller's array of those.
struct RangeOut {
  uint64_t count = 0, first = 0, last = 0;
  uint32_t hist[kQueryBins] = {0};
  uint32_t pad = 0;
};

struct KernelTimes {
  double sieve_ms = 0, stitch_ms = 0, rescan_ms = 0, total_ms = 0;
  uint64_t segments = 0;
};

// Device buffer (RAII).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n);
  ~DevBuf();
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept;
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// One captured batch pipeline (see Engine::batch_ranges).
struct BatchPlan {
  uint32_t qcap = 0, scap = 0;
  uint32_t seg_bytes = 0, tier1_p = 0;
  int block = 0, wgs_per_cu = 0;
  uint64_t table_limit = 0;
  const uint8_t* cache = nullptr;  // resident sieved bitmap of [7, cache_hi] (null: none)
  uint64_t cache_hi = 0;
  size_t in_bytes = 0, out_bytes = 0;
  void* h_in = nullptr;   // pinned: header | queries | segment->query map
  void* h_out = nullptr;  // pinned: count | first | last | histograms
  DevBuf d_in, d_sum, d_out, d_work;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  BatchPlan() = default;
  ~BatchPlan();
  BatchPlan(const BatchPlan&) = delete;
  BatchPlan& operator=(const BatchPlan&) = delete;
};

using ProgressFn = std::function<void(int seg, int total, uint64_t in_seg, uint64_t cumulative)>;

// Prefix counts of sweep segments (segments of seg_bytes from k = 0):
// cum[s] = number of primes below 30 * s * seg_bytes, cum[nseg] = pi(hi).
struct NthIndex {
  uint64_t hi = 0;
  uint32_t seg_bytes = 0;
  std::vector<uint64_t> cum;
};

// The no-allocation count (Engine::count_range_noalloc) keeps its work
// counter, partial count and per-segment counts in a fixed device array of
// kNoAllocHeadBytes + kNoAllocMaxSegs * sizeof(SegSummary) bytes (about
// 256 KiB), so it serves ranges of at most kNoAllocMaxSegs segments.
constexpr uint32_t kNoAllocMaxSegs = 16384;
constexpr size_t kNoAllocHeadBytes = 256;

class Engine {
 public:
  static Engine& get();  // initialises the device on first use; throws if none
  static Engine* peek();  // the engine if get() has succeeded before, else null

  // kind 0: sweeps (prime_count, full statistics, bitmaps); kind 1: batched ranges
  const TuneConfig& config(int kind = 0) const { return cfgs_[kind & 1]; }
  void set_config(const TuneConfig& c, int kind = -1);  // -1: both kinds
  int num_cus() const { return ncu_; }
  std::string device_name() const { return name_; }
  hipStream_t stream() const { return stream_; }

  // pi(N) counting primes <= N (N >= 2).
  uint64_t prime_count(uint64_t N);
  // Count primes in [lo, hi] (lo >= 7 handled by the bitmap; smaller lo adds seeds).
  uint64_t count_range(uint64_t lo, uint64_t hi);
  // Full statistics over [2, N].
  SweepStats sweep_stats(uint64_t N, uint64_t cp_interval, uint32_t large_threshold,
                         const ProgressFn& progress);
  // Batched ranges: one persistent launch over the flattened segment list.
  void batch_ranges(const uint64_t* a, const uint64_t* b, size_t n, RangeOut* out);
  // All primes in [a, b], ascending (true primes; 2, 3, 5 included).  Sets
  // *overflow (and returns early) when there are more than max_count.
  std::vector<uint64_t> primes_in_range(uint64_t a, uint64_t b, uint64_t max_count, bool* overflow);
  // Sieve [max(lo,7), hi] into a device bitmap (byte k - k_base, bits outside
  // the range cleared).  Returns the bitmap length in bytes.
  uint64_t bitmap_range(uint64_t lo, uint64_t hi, DevBuf& out, uint64_t* k_base);

  // Callers around the sieve (engine_callers.hip)
  // Goldbach counterexamples in [4, N], ascending, the first max_ce; partners
  // limited to p <= pmax (0: none), first-pass scan to scan_limit (0: the
  // table); *total (if given) receives how many there are in all
  std::vector<uint64_t> goldbach_upto(uint64_t N, uint64_t max_ce, uint64_t pmax = 0, uint64_t scan_limit = 0,
                                      uint64_t* total = nullptr);
  std::vector<uint64_t> goldbach_counts(const uint64_t* ns, size_t k);    // partition counts
  // partition counts (p1 <= plim) over a bitmap of [0, >= nmax] built by bitmap_range
  std::vector<uint64_t> goldbach_counts_bm(const DevBuf& bm, const uint64_t* ns, size_t k, uint64_t nmax,
                                           uint64_t plim);
  uint64_t sophie_germain(uint64_t N);                                    // p <= N, 2p+1 prime
  void is_prime_batch(const uint64_t* x, size_t n, uint8_t* out);
  void next_prev_batch(const uint64_t* x, size_t n, uint64_t* out, int dir);
  uint64_t nth_prime(uint64_t n);                                         // 1-indexed
  // Prefix-count index covering [2, hi] for the sweep configuration c; kept
  // and reused while a later request fits in it (same segment size).
  std::shared_ptr<const NthIndex> nth_index(uint64_t hi, const TuneConfig& c);

  KernelTimes last_times() const { return times_; }

  // LDS rate probe (launch_lds_probe): lane-operations per clock per CU at
  // the clock measured during the run, that clock, and the kernel time.
  struct LdsProbe {
    double ops_per_clk_cu = 0, clock_hz = 0, ms = 0;
  };
  LdsProbe lds_probe(int kind, uint32_t iters);

  // Prime count of [a, b] written to d_out[0] (device memory) by kernels
  // enqueued on stream st, with no allocation, no copy and no wait: every
  // buffer beyond d_out is the fixed scratch array.  Calls that overlap on
  // the device race on that array.  Ranges of more than kNoAllocMaxSegs
  // segments are rejected; the current device must be the engine's.
  void count_range_noalloc(uint64_t a, uint64_t b, unsigned long long* d_out, hipStream_t st);

  // ---- batched ranges split into host layout, staging and read-back, so a
  // caller with its own stream (the request server) can run them
  // asynchronously.  All three need only the resident tables.
  struct BatchLayout {
    std::vector<uint32_t> order;  // caller index per device query, costliest first
    uint32_t nq = 0, nseg = 0;
  };
  BatchLayout layout_batch(const uint64_t* a, const uint64_t* b, size_t n, uint32_t seg_bytes) const;
  // capture a batch pipeline for up to qcap queries / scap segments on stream st
  // queries with b <= cache_hi read their bytes from cache (a sieved bitmap of
  // [7, cache_hi] from byte 0, as bitmap_range writes it) instead of sieving
  std::unique_ptr<BatchPlan> make_batch_plan(uint32_t qcap, uint32_t scap, const TuneConfig& c, hipStream_t st,
                                             const uint8_t* cache = nullptr, uint64_t cache_hi = 0);
  static void stage_batch(BatchPlan& P, const BatchLayout& L, const uint64_t* a, const uint64_t* b);
  // results (including the seeds 2, 3, 5) into out[0..n); P may be null when L.nq == 0
  static void read_batch(const BatchPlan* P, const BatchLayout& L, const uint64_t* a, const uint64_t* b, size_t n,
                         RangeOut* out);
  SieveArgs base_args(uint32_t nseg, const TuneConfig& c) const;
  int grid_for(uint32_t nseg, const TuneConfig& c) const;
  // per-segment prime counts of [7, hi] (segments of c.seg_bytes from k = 0)
  uint32_t segment_counts(uint64_t hi, std::vector<uint32_t>& counts, const TuneConfig& c);

  // exposed for tests and tuning
  void ensure_tables(uint64_t max_value);
  uint32_t tier_index(uint32_t p) const;  // #sieving primes < p

  std::mutex mu;  // one computation at a time per process

 private:
  Engine();
  ~Engine();
  SieveArgs base_args(uint32_t nseg) const { return base_args(nseg, cfg_); }
  int grid_for(uint32_t nseg) const { return grid_for(nseg, cfg_); }
  void load_tune_file();
  uint64_t count_bitmap(uint64_t lo, uint64_t hi);  // wheel-bitmap primes in [max(lo,7), hi]

  BatchPlan& batch_plan(uint32_t nq, uint32_t nseg);
  std::vector<std::unique_ptr<BatchPlan>> plans_;

  // reusable device workspace, grown on demand
  enum WsSlot {
    kWsWork, kWsTotals, kWsSummaries, kWsHist, kWsBuckets, kWsLarge, kWsRec, kWsRescan,
    kWsRescanOut, kWsCounters, kWsQueries, kWsSegQuery, kWsQHist, kWsQCount, kWsList, kWsWaveGaps, kWsStage, kWsCount_
  };
  DevBuf& ws(int slot, size_t bytes);
  std::vector<DevBuf> ws_;
  std::shared_ptr<const NthIndex> nth_cache_;
  // persistent pinned host staging for downloads, grown on demand
  void* pinned(size_t bytes);
  void* pinned_ = nullptr;
  size_t pinned_bytes_ = 0;

  int dev_ = 0, ncu_ = 0;
  std::string name_;
  hipStream_t stream_ = nullptr;
  TuneConfig cfgs_[2];
  TuneConfig cfg_;  // the configuration of the call in flight
  // resident tables
  std::vector<uint32_t> h_primes_;  // sieving primes >= kFirstSievingPrime (plain values)
  uint32_t* d_primes_ = nullptr;
  uint32_t* d_pitab_ = nullptr;
  uint32_t pitab_len_ = 0;
  uint32_t* d_presieve_ = nullptr;
  uint32_t presieve_off_[kPresieveGroups] = {0};
  uint32_t first_word_ = 0;
  uint64_t table_limit_ = 0;  // tables cover sieving primes <= table_limit_
  uint32_t tier0_p_ = kTier0P;  // primes below: kernel tier 0
  uint32_t presieve_copy_words_[kPresieveGroups] = {0};
  uint32_t tier2_div_ = 5;      // primes below seg_bytes / this: kernel tier 2
  KernelTimes times_;
  // device address of the no-allocation scratch array (null if the lookup
  // at start-up failed; scratch_err_ says why)
  void* scratch_ = nullptr;
  std::string scratch_err_;
};

void hip_check(hipError_t e, const char* what);
// Full statistics for N < 7 (no device needed).
SweepStats small_sweep_stats(uint64_t N);

}  // namespace voss
