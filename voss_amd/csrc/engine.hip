// engine.hip — host orchestration of the gfx950 sieve.
//
// Host-side flow of a full-statistics sweep over [2, N] (the reference's
// compute_full_impl, library/src/primes/context.cu:212-559):
//   1. one persistent launch of the segment kernel over every LDS segment
//      (in place of one graph of ~27k per-prime launches per 5e10 segment,
//      an extract, a thrust sort, a mod-4 kernel and a gap kernel);
//   2. stitch_kernel joins the segments in order (boundary gaps, records);
//   3. rescan_kernel lists the records inside the few flagged segments;
//   4. a handful of small device-to-host copies.
// The seeds 2, 3, 5 and the reference's fix-ups (hist[1] += 1, hist[2] += 1,
// pi(x;4,1) += 1 for 5, pi(x;4,3) += 1 for 3; context.cu:513-524) are applied
// here on the host.
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <atomic>
#include <cstring>
#include <dlfcn.h>

#include "kernels.h"

namespace voss {

namespace {
const uint32_t kW30[8] = {1, 7, 11, 13, 17, 19, 23, 29};
constexpr int kErrOom = 4, kErrDevice = 5, kErrArg = 6, kErrInternal = 7;

uint64_t isqrt64(uint64_t n) {
  uint64_t r = (uint64_t)std::sqrt((double)n);
  while (r * r > n) --r;
  while ((r + 1) * (r + 1) <= n) ++r;
  return r;
}

uint32_t env_u32(const char* name, uint32_t dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return (uint32_t)std::strtoul(v, nullptr, 0);
}

// Timing events come from a per-thread free list: creating and destroying
// four per call cost more than the calls' host work.
struct Event {
  hipEvent_t e = nullptr;
  static std::vector<hipEvent_t>& pool() {
    static thread_local std::vector<hipEvent_t> p;
    return p;
  }
  Event() {
    auto& p = pool();
    if (!p.empty()) { e = p.back(); p.pop_back(); }
    else hip_check(hipEventCreate(&e), "hipEventCreate");
  }
  ~Event() { if (e) pool().push_back(e); }
};
float elapsed(const Event& a, const Event& b) {
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, a.e, b.e), "hipEventElapsedTime");
  return ms;
}
}  // namespace

// Scratch of Engine::count_range_noalloc, the one entry point that never
// allocates: bytes [0, 4) the segment work counter, [8, 16) the kernel's
// running total, then one SegSummary per segment (kNoAllocMaxSegs at most).
// Each GPU holds its own copy; host code uses only the address that
// hipGetSymbolAddress returns for it (Engine::scratch_), never this name.
static __device__ __align__(256) uint8_t scratch[kNoAllocHeadBytes + (size_t)kNoAllocMaxSegs * sizeof(SegSummary)];

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(kErrDevice, std::string(what) + " failed: " + hipGetErrorString(e));
}

DevBuf::DevBuf(size_t n) : bytes(n) {
  if (!n) return;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    throw Error(kErrOom, "hipMalloc(" + std::to_string(n) + " bytes) failed: " + hipGetErrorString(e));
  }
}
DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}
DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
  if (this != &o) {
    if (p) (void)hipFree(p);
    p = o.p;
    bytes = o.bytes;
    o.p = nullptr;
    o.bytes = 0;
  }
  return *this;
}

namespace {
std::atomic<Engine*> g_engine{nullptr};
}

Engine& Engine::get() {
  static Engine* inst = new Engine();  // lives until process exit
  g_engine.store(inst, std::memory_order_release);
  return *inst;
}

Engine* Engine::peek() { return g_engine.load(std::memory_order_acquire); }

Engine::Engine() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    // the reference throws here and its entry points map that to
    // VOSS_ERR_CUDA (prime_count.cu:147-151, 316-318)
    (void)hipGetLastError();
    throw Error(kErrDevice, "no HIP device available");
  }
  hip_check(hipGetDevice(&dev_), "hipGetDevice");
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, dev_), "hipGetDeviceProperties");
  ncu_ = prop.multiProcessorCount;
  name_ = prop.gcnArchName;
  hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  ws_.resize(kWsCount_);

  load_tune_file();
  tier0_p_ = env_u32("VOSS_TIER0_P", kTier0P);
  tier2_div_ = std::max<uint32_t>(1, env_u32("VOSS_TIER2_DIV", 5));
  if (std::getenv("VOSS_SEG_BYTES") || std::getenv("VOSS_BLOCK") || std::getenv("VOSS_WGS_PER_CU") ||
      std::getenv("VOSS_TIER1_P")) {
    TuneConfig c = cfgs_[0];
    c.seg_bytes = env_u32("VOSS_SEG_BYTES", c.seg_bytes);
    c.block = (int)env_u32("VOSS_BLOCK", (uint32_t)c.block);
    c.wgs_per_cu = (int)env_u32("VOSS_WGS_PER_CU", (uint32_t)c.wgs_per_cu);
    c.tier1_p = env_u32("VOSS_TIER1_P", c.tier1_p);
    set_config(c);
  }
  cfg_ = cfgs_[0];

  // presieve patterns: in pattern g, byte j has bit o set iff 30j + W30[o] has
  // no factor among the group's primes; each is stored for one period plus a
  // full segment so a segment never wraps.  Pattern 0 is a segment's initial
  // bytes, the others (the read patterns) are ANDed in as it is read.  Each
  // is built as a small sieve: the multiples of q in class o are the bytes
  // j = j0 (mod q) with 30 j0 + W30[o] = 0 (mod q).
  static const uint32_t groups[kPresieveGroups][5] = {
      {7, 11, 13, 17, 19}, {23, 29, 31, 0, 0},  {37, 41, 43, 0, 0},   {47, 53, 59, 0, 0},
      {61, 67, 71, 0, 0},  {73, 79, 83, 0, 0},  {89, 97, 101, 0, 0}, {103, 107, 109, 0, 0}};
  std::vector<uint32_t> words;
  for (int gi = 0; gi < kPresieveGroups; ++gi) {
    presieve_off_[gi] = (uint32_t)words.size();
    const size_t len = ((size_t)kPresievePeriods[gi] + kMaxSegBytes + 4096 + 3) & ~(size_t)3;
    std::vector<uint8_t> pat(len + 4, 0xFF);
    for (int qi = 0; qi < 5 && groups[gi][qi]; ++qi) {
      const uint32_t q = groups[gi][qi];
      for (int o = 0; o < 8; ++o) {
        uint32_t j0 = 0;
        while ((30ull * j0 + kW30[o]) % q != 0) ++j0;  // j0 < q
        const uint8_t clear = (uint8_t)~(1u << o);
        for (size_t j = j0; j < len + 4; j += q) pat[j] &= clear;
      }
    }
    // four copies, copy c starting c bytes in, so a read at any byte phase
    // starts on a word of one of them (the kernel copies it to LDS directly)
    presieve_copy_words_[gi] = (uint32_t)(len / 4);
    for (int c = 0; c < 4; ++c) {
      const size_t w0 = words.size();
      words.resize(w0 + len / 4);
      std::memcpy(words.data() + w0, pat.data() + c, len);
    }
  }
  hip_check(hipMalloc(&d_presieve_, words.size() * 4), "hipMalloc(presieve)");
  hip_check(hipMemcpy(d_presieve_, words.data(), words.size() * 4, hipMemcpyHostToDevice), "hipMemcpy(presieve)");
  // true prime bitmap of bytes 0..3 (values below 120)
  first_word_ = 0;
  for (uint32_t byte = 0; byte < 4; ++byte)
    for (uint32_t o = 0; o < 8; ++o) {
      const uint32_t v = 30 * byte + kW30[o];
      bool pr = v >= 2;
      for (uint32_t d = 2; d * d <= v; ++d) if (v % d == 0) { pr = false; break; }
      if (pr) first_word_ |= 1u << (8 * byte + o);
    }
  hip_check(prepare_kernels(lds_bytes_for(kMaxSegBytes)), "prepare_kernels");
  // sieving primes up to 2^24 cover every value below 2.8e14; built once so
  // concurrent users (the request server's slots) never see them move
  ensure_tables(1ull << 48);
  // the device address of this GPU's scratch array (the engine serves one
  // device, current here); a failed lookup disables only the no-allocation
  // entry point
  {
    void* p = nullptr;
    const hipError_t se = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (se == hipSuccess && p) {
      scratch_ = p;
    } else {
      (void)hipGetLastError();
      scratch_err_ = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(se);
    }
  }
}

Engine::~Engine() {}

// Tuning file written by the autotuner (voss_amd/autotune.py): lines
// "<sweep|batch> <seg_bytes> <block> <wgs_per_cu> <tier1_p>".  Looked up in
// $VOSS_TUNE_FILE, else tuned.cfg next to this library.
void Engine::load_tune_file() {
  std::string path;
  if (const char* env = std::getenv("VOSS_TUNE_FILE")) {
    path = env;
  } else {
    Dl_info info;
    if (dladdr((void*)&hip_check, &info) && info.dli_fname) {
      path = info.dli_fname;
      const size_t slash = path.rfind('/');
      path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuned.cfg";
    }
  }
  FILE* f = path.empty() ? nullptr : std::fopen(path.c_str(), "r");
  if (!f) return;
  char kind[32];
  unsigned seg = 0, t1 = 0;
  int blk = 0, wgs = 0;
  while (std::fscanf(f, "%31s %u %d %d %u", kind, &seg, &blk, &wgs, &t1) == 5) {
    TuneConfig c;
    c.seg_bytes = seg;
    c.block = blk;
    c.wgs_per_cu = wgs;
    c.tier1_p = t1;
    try {
      if (std::string(kind) == "sweep") set_config(c, 0);
      else if (std::string(kind) == "batch") set_config(c, 1);
    } catch (const Error&) {
      // an invalid line keeps the built-in default
    }
  }
  std::fclose(f);
}

void Engine::set_config(const TuneConfig& c, int kind) {
  if (c.block != 512 && c.block != 1024) throw Error(kErrArg, "block must be 512 or 1024");
  const uint32_t nw = (uint32_t)c.block / 64;
  if (c.seg_bytes == 0 || c.seg_bytes > kMaxSegBytes || c.seg_bytes % (1024 * nw) != 0)
    throw Error(kErrArg, "seg_bytes must be a multiple of 1 KiB x waves and at most 144 KiB");
  if (c.wgs_per_cu < 1 || c.wgs_per_cu > 8) throw Error(kErrArg, "wgs_per_cu must be in [1, 8]");
  if ((uint64_t)lds_bytes_for(c.seg_bytes) * (uint64_t)c.wgs_per_cu > 160 * 1024)
    throw Error(kErrArg, "segment LDS x workgroups per CU exceeds 160 KiB");
  if (kind < 0 || kind == 0) cfgs_[0] = c;
  if (kind < 0 || kind == 1) cfgs_[1] = c;
}

void Engine::ensure_tables(uint64_t max_value) {
  const uint64_t need = isqrt64(max_value) + 2;
  if (need <= table_limit_) return;
  // the table is built once, at its largest, in the constructor: it never
  // moves under a caller that holds its address
  if (table_limit_) throw Error(kErrArg, "value beyond the sieving-prime table (2.8e14)");
  uint64_t lim = std::max<uint64_t>(need + need / 4, 1u << 16);
  if (lim > (1ull << 24) - 1) lim = (1ull << 24) - 1;  // 1.67e7 > sqrt(1e14)
  std::vector<uint8_t> s(lim + 1, 1);
  s[0] = s[1] = 0;
  for (uint64_t i = 2; i * i <= lim; ++i)
    if (s[i]) for (uint64_t j = i * i; j <= lim; j += i) s[j] = 0;
  std::vector<uint32_t> entries;
  h_primes_.clear();
  for (uint64_t p = kFirstSievingPrime; p <= lim; ++p) {
    if (!s[p]) continue;
    uint32_t ridx = 0;
    for (uint32_t o = 0; o < 8; ++o) if (kW30[o] == p % 30) ridx = o;
    h_primes_.push_back((uint32_t)p);
    entries.push_back((uint32_t)p | (ridx << 24));
  }
  const uint32_t len = (uint32_t)(lim / kPiTabStep + 2);
  std::vector<uint32_t> pitab(len);
  size_t k = 0;
  for (uint32_t i = 0; i < len; ++i) {
    const uint64_t bound = (uint64_t)i * kPiTabStep;
    while (k < h_primes_.size() && h_primes_[k] <= bound) ++k;
    pitab[i] = (uint32_t)k;
  }
  if (d_primes_) (void)hipFree(d_primes_);
  if (d_pitab_) (void)hipFree(d_pitab_);
  d_primes_ = nullptr;
  d_pitab_ = nullptr;
  hip_check(hipMalloc(&d_primes_, entries.size() * 4 + 4), "hipMalloc(primes)");
  hip_check(hipMalloc(&d_pitab_, (size_t)len * 4), "hipMalloc(pi_tab)");
  hip_check(hipMemcpy(d_primes_, entries.data(), entries.size() * 4, hipMemcpyHostToDevice), "hipMemcpy(primes)");
  hip_check(hipMemcpy(d_pitab_, pitab.data(), (size_t)len * 4, hipMemcpyHostToDevice), "hipMemcpy(pi_tab)");
  pitab_len_ = len;
  table_limit_ = lim;
}

uint32_t Engine::tier_index(uint32_t p) const {
  return (uint32_t)(std::lower_bound(h_primes_.begin(), h_primes_.end(), p) - h_primes_.begin());
}

SieveArgs Engine::base_args(uint32_t nseg, const TuneConfig& c) const {
  SieveArgs A;
  std::memset(&A, 0, sizeof(A));
  A.primes = d_primes_;
  A.pi_tab = d_pitab_;
  A.pi_tab_len = pitab_len_;
  A.presieve = d_presieve_;
  for (int k = 0; k < kPresieveGroups; ++k) A.presieve_off[k] = presieve_off_[k];
  for (int k = 0; k < kPresieveGroups; ++k) A.presieve_copy_words[k] = presieve_copy_words_[k];
  A.first_word = first_word_;
  A.seg_bytes = c.seg_bytes;
  A.nseg = nseg;
  A.tier0_end = tier_index(tier0_p_);
  A.tier1_end = tier_index(c.tier1_p);
  A.tier2_end = tier_index(c.seg_bytes / tier2_div_);
  A.tier3_end = tier_index(c.seg_bytes);
  A.large_threshold = 0xFFFFFFFFu;
  return A;
}

int Engine::grid_for(uint32_t nseg, const TuneConfig& c) const {
  uint64_t g = (uint64_t)ncu_ * (uint64_t)c.wgs_per_cu;
  if (g > nseg) g = nseg;
  return (int)std::max<uint64_t>(g, 1);
}

void* Engine::pinned(size_t bytes) {
  if (bytes > pinned_bytes_) {
    if (pinned_) (void)hipHostFree(pinned_);
    pinned_ = nullptr;
    pinned_bytes_ = 0;
    const size_t want = std::max(bytes, (size_t)1 << 20);
    hip_check(hipHostMalloc(&pinned_, want, hipHostMallocDefault), "hipHostMalloc(staging)");
    pinned_bytes_ = want;
  }
  return pinned_;
}

DevBuf& Engine::ws(int slot, size_t bytes) {
  // ws_ is sized once in the constructor: references handed out stay valid
  if (slot < 0 || slot >= (int)ws_.size()) throw Error(7, "bad workspace slot");
  if (ws_[slot].bytes < bytes) {
    ws_[slot] = DevBuf();
    ws_[slot] = DevBuf(std::max(bytes, (size_t)256));
  }
  return ws_[slot];
}

// ------------------------------------------------------------------ counting
uint64_t Engine::count_bitmap(uint64_t lo, uint64_t hi) {
  cfg_ = cfgs_[0];
  if (lo < 7) lo = 7;
  if (hi < lo) return 0;
  ensure_tables(hi);
  const uint32_t S = cfg_.seg_bytes;
  const uint64_t k_base = lo / 30, k_end = hi / 30 + 1;
  const uint64_t nseg64 = (k_end - k_base + S - 1) / S;
  if (nseg64 > 0xFFFFFFF0ull) throw Error(kErrArg, "range too large");
  const uint32_t nseg = (uint32_t)nseg64;
  DevBuf& wk = ws(kWsWork, 64);
  DevBuf& tot = ws(kWsTotals, 64);
  hip_check(hipMemsetAsync(wk.p, 0, 64, stream_), "memset");
  hip_check(hipMemsetAsync(tot.p, 0, 64, stream_), "memset");
  SieveArgs A = base_args(nseg);
  A.k_base = k_base;
  A.k_end = k_end;
  A.lo_val = lo;
  A.hi_val = hi;
  A.work = wk.as<unsigned int>();
  A.totals = tot.as<unsigned long long>();
  Event e0, e1;
  hip_check(hipEventRecord(e0.e, stream_), "hipEventRecord");
  hip_check(launch_sieve(kModeCount, cfg_.block, A, grid_for(nseg), stream_), "launch(count)");
  hip_check(hipEventRecord(e1.e, stream_), "hipEventRecord");
  unsigned long long c = 0;
  hip_check(hipMemcpyAsync(&c, tot.p, 8, hipMemcpyDeviceToHost, stream_), "hipMemcpy(count)");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  times_ = KernelTimes();
  times_.sieve_ms = elapsed(e0, e1);
  times_.total_ms = times_.sieve_ms;
  times_.segments = nseg;
  return c;
}

void Engine::count_range_noalloc(uint64_t a, uint64_t b, unsigned long long* d_out, hipStream_t st) {
  if (!scratch_) throw Error(kErrInternal, scratch_err_);
  int cur = -1;
  hip_check(hipGetDevice(&cur), "hipGetDevice");
  if (cur != dev_) throw Error(kErrArg, "the current device is not the one the library was initialised on");
  const TuneConfig c = cfgs_[0];
  const uint32_t S = c.seg_bytes;
  uint64_t seeds = 0;
  for (uint64_t sd : {2ull, 3ull, 5ull}) seeds += (sd >= a && sd <= b) ? 1 : 0;
  const uint64_t lo = std::max<uint64_t>(a, 7);
  uint8_t* base = static_cast<uint8_t*>(scratch_);
  SegSummary* sums = reinterpret_cast<SegSummary*>(base + kNoAllocHeadBytes);
  uint32_t nseg = 0;
  if (b >= lo) {
    const uint64_t k_base = lo / 30, k_end = b / 30 + 1;
    const uint64_t nseg64 = (k_end - k_base + S - 1) / S;
    if (nseg64 > kNoAllocMaxSegs)
      throw Error(2, "range needs " + std::to_string(nseg64) + " segments; the no-allocation count holds at most " +
                         std::to_string(kNoAllocMaxSegs) + " (" + std::to_string(30ull * S * kNoAllocMaxSegs) +
                         " integers)");
    nseg = (uint32_t)nseg64;
    // the work counter and running total start at zero (a fill kernel: a
    // memset node would not replay correctly inside a captured graph)
    hip_check(launch_zero(reinterpret_cast<uint32_t*>(base), kNoAllocHeadBytes / 4, st), "launch(zero)");
    SieveArgs A = base_args(nseg, c);
    A.k_base = k_base;
    A.k_end = k_end;
    A.lo_val = lo;
    A.hi_val = b;
    A.work = reinterpret_cast<unsigned int*>(base);
    A.totals = reinterpret_cast<unsigned long long*>(base + 8);
    A.summaries = sums;  // every segment's count is written: nothing to clear
    hip_check(launch_sieve(kModeCount, c.block, A, grid_for(nseg, c), st), "launch(count)");
  }
  hip_check(launch_sum_counts(sums, nseg, seeds, d_out, st), "launch(sum_counts)");
}

Engine::LdsProbe Engine::lds_probe(int kind, uint32_t iters) {
  if (kind < 0 || kind > 3) throw Error(kErrArg, "probe kind must be 0..3");
  if (iters == 0 || iters > (1u << 24)) throw Error(kErrArg, "probe iterations must be in [1, 2^24]");
  DevBuf d(64);
  hip_check(hipMemsetAsync(d.p, 0, 64, stream_), "memset");
  Event e0, e1;
  hip_check(hipEventRecord(e0.e, stream_), "hipEventRecord");
  hip_check(launch_lds_probe(kind, (uint32_t)ncu_, iters, d.as<unsigned long long>(), d.as<uint32_t>() + 8, stream_),
            "launch(lds_probe)");
  hip_check(hipEventRecord(e1.e, stream_), "hipEventRecord");
  unsigned long long clk[2] = {0, 0};
  hip_check(hipMemcpyAsync(clk, d.p, 16, hipMemcpyDeviceToHost, stream_), "D2H");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  LdsProbe R;
  R.ms = elapsed(e0, e1);
  R.clock_hz = clk[1] ? (double)clk[0] / ((double)clk[1] / 100e6) : 0.0;  // real-time counter: 100 MHz
  const double ops = (double)ncu_ * 1024.0 * 8.0 * (double)iters;
  if (R.ms > 0 && R.clock_hz > 0) R.ops_per_clk_cu = ops / (R.ms / 1e3) / (double)ncu_ / R.clock_hz;
  return R;
}

uint64_t Engine::count_range(uint64_t lo, uint64_t hi) {
  if (hi < lo) return 0;
  uint64_t c = 0;
  for (uint64_t sd : {2ull, 3ull, 5ull}) c += (sd >= lo && sd <= hi) ? 1 : 0;
  return c + count_bitmap(lo, hi);
}

uint64_t Engine::prime_count(uint64_t N) {
  if (N < 2) return 0;
  if (N < 7) return N < 3 ? 1 : (N < 5 ? 2 : 3);
  return 3 + count_bitmap(7, N);
}

// ------------------------------------------------------------------ full statistics
// N < 7: the reference short-circuits (context.cu:228-247) with pi = 1/2/3,
// hist[1] / hist[2] set and the mod-4 tallies left at 0.
SweepStats small_sweep_stats(uint64_t N) {
  SweepStats R;
  R.hist.assign(100000, 0);
  const uint64_t seeds[3] = {2, 3, 5};
  uint64_t nseed = 0;
  for (uint64_t sd : seeds) nseed += sd <= N ? 1 : 0;
  R.prime_count = nseed;
  if (N >= 3) { R.hist[1] = 1; R.records.push_back({3, 1}); }
  if (N >= 5) { R.hist[2] = 1; R.records.push_back({5, 2}); }
  R.last_prime = nseed ? seeds[nseed - 1] : 0;
  R.max_gap = nseed >= 3 ? 2 : (nseed == 2 ? 1 : 0);
  R.pi_3_2 = N >= 5 ? 2 : (N >= 2 ? 1 : 0);
  return R;
}

SweepStats Engine::sweep_stats(uint64_t N, uint64_t cp_interval, uint32_t large_threshold,
                               const ProgressFn& progress) {
  cfg_ = cfgs_[0];
  SweepStats R;
  if (N < 2) throw Error(1, "N must be >= 2");
  const uint32_t S = cfg_.seg_bytes;
  if (cp_interval && cp_interval < 30ull * S)
    throw Error(kErrArg, "checkpoint interval must be at least " + std::to_string(30ull * S));
  const uint64_t ncp = cp_interval ? N / cp_interval : 0;  // checkpoints reported to the caller
  if (N < 7) return small_sweep_stats(N);
  const uint64_t nseed = 3;
  R.records.push_back({3, 1});
  R.records.push_back({5, 2});
  ensure_tables(N);
  const uint64_t k_end = N / 30 + 1;
  const uint64_t nseg64 = (k_end + S - 1) / S;
  if (nseg64 > 0x7FFFFFF0ull) throw Error(kErrArg, "N too large for the segment size");
  const uint32_t nseg = (uint32_t)nseg64;
  // Record lists cannot overflow: every boundary record, every listed wave
  // range and every rescanned record strictly raises a running maximum that
  // the stitch's block scan carries across its threads, and the listed
  // values are 16-bit (gaps and per-wave bounds below 65536), so each list
  // holds fewer than 2^16 entries (1 MiB each).
  const uint32_t large_cap = 1u << 20, rec_cap = 1u << 16, rescan_cap = 1u << 16, rout_cap = 1u << 16;

  // Progress reports per reference segment (context.cu:260-269, 433-436):
  // SEG_NUM = min(ceil(N/30)*30 + 10, 5e10 + 10) integers, NUM_SEG =
  // ceil(N / SEG_NUM) calls of (seg, NUM_SEG, primes in seg, cumulative).
  // The kernel segments stay on their uniform grid: the checkpoint buckets,
  // set to SEG_NUM, count each reference segment exactly (a kernel segment
  // that straddles a boundary splits its tallies), and launch L ends with
  // the kernel segment that holds reference segment L's last byte, so
  // bucket L is final when launch L completes.
  uint64_t ref_seg = 0, nref = 1;
  if (progress) {
    if (cp_interval) throw Error(kErrArg, "a progress callback cannot be combined with checkpoints");
    constexpr uint64_t kMaxRefSeg = 50000000010ull;
    ref_seg = N < kMaxRefSeg ? ((N + 29) / 30) * 30 + 10 : kMaxRefSeg;
    nref = (N + ref_seg - 1) / ref_seg;
    if (nref > 1) cp_interval = ref_seg;
  }
  const uint32_t nlaunch = (uint32_t)nref;
  // kernel segments [launch_end[L-1], launch_end[L]) go to launch L
  std::vector<uint32_t> launch_end(nlaunch, nseg);
  for (uint32_t L = 0; L + 1 < nlaunch; ++L) {
    const uint64_t bytes = (uint64_t)(L + 1) * (ref_seg / 30);  // reference segments 0..L
    launch_end[L] = (uint32_t)std::min<uint64_t>(nseg, (bytes + S - 1) / S);
  }
  const uint64_t nbuckets = cp_interval ? (N + cp_interval - 1) / cp_interval : 0;

  const uint32_t waves = (uint32_t)cfg_.block / 64;
  DevBuf& d_sum = ws(kWsSummaries, (size_t)nseg * sizeof(SegSummary));
  DevBuf& d_wg = ws(kWsWaveGaps, (size_t)nseg * waves * 4);
  DevBuf& d_tot = ws(kWsTotals, 64);
  DevBuf& d_hist = ws(kWsHist, (size_t)kGapBins * 8);
  DevBuf& d_bk = ws(kWsBuckets, (size_t)std::max<uint64_t>(nbuckets, 1) * 3 * 8);
  DevBuf& d_large = ws(kWsLarge, (size_t)large_cap * sizeof(GapRec));
  DevBuf& d_rec = ws(kWsRec, (size_t)rec_cap * sizeof(GapRec));
  DevBuf& d_resc = ws(kWsRescan, (size_t)rescan_cap * sizeof(RescanItem));
  DevBuf& d_rout = ws(kWsRescanOut, (size_t)rout_cap * sizeof(GapRec));
  DevBuf& d_cnt = ws(kWsCounters, 64);  // [0] n_large [1] n_rec [2] n_rescan [3] n_rout, [8..] u64 final
  DevBuf& d_work = ws(kWsWork, (size_t)std::max<uint32_t>(nlaunch, 1) * 4 + 64);

  {
    // every accumulator cleared by one launch
    SpanList Z;
    std::memset(&Z, 0, sizeof(Z));
    auto zero = [&](DevBuf& b, size_t bytes) { Z.s[Z.n++] = CopySpan{nullptr, b.as<uint32_t>(), bytes / 4}; };
    zero(d_tot, 64);
    zero(d_hist, (size_t)kGapBins * 8);
    zero(d_bk, (size_t)std::max<uint64_t>(nbuckets, 1) * 3 * 8);
    zero(d_cnt, 64);
    zero(d_work, (size_t)std::max<uint32_t>(nlaunch, 1) * 4 + 64);
    hip_check(launch_spans(Z, stream_), "launch(zero)");
  }

  SieveArgs A = base_args(nseg);
  A.k_base = 0;
  A.k_end = k_end;
  A.lo_val = 7;
  A.hi_val = N;
  A.totals = d_tot.as<unsigned long long>();
  A.summaries = d_sum.as<SegSummary>();
  A.wave_gaps = d_wg.as<uint32_t>();
  A.hist = d_hist.as<unsigned long long>();
  A.buckets = d_bk.as<unsigned long long>();
  A.cp_interval = cp_interval;
  A.nbuckets = (uint32_t)nbuckets;
  A.large = d_large.as<GapRec>();
  A.n_large = d_cnt.as<unsigned int>() + 0;
  A.large_cap = large_cap;
  A.large_threshold = large_threshold;

  Event e0, e1, e2, e3;
  hip_check(hipEventRecord(e0.e, stream_), "hipEventRecord");
  uint64_t cumulative = nseed;
  for (uint32_t L = 0; L < nlaunch; ++L) {
    SieveArgs AL = A;
    AL.seg_base = L ? launch_end[L - 1] : 0;
    AL.nseg = launch_end[L] - AL.seg_base;
    AL.work = d_work.as<unsigned int>() + L;
    // a reference segment wholly inside the previous launch's last kernel
    // segment (N just past a boundary) needs no launch of its own
    if (AL.nseg)
      hip_check(launch_sieve(kModeStats, cfg_.block, AL, grid_for(AL.nseg), stream_), "launch(stats)");
    if (progress) {
      // primes of reference segment L: its bucket, or every prime found so far
      unsigned long long c = 0;
      const void* src = nref > 1 ? (const void*)(d_bk.as<unsigned long long>() + 3 * L) : d_tot.p;
      hip_check(hipMemcpyAsync(&c, src, 8, hipMemcpyDeviceToHost, stream_), "hipMemcpy(progress)");
      hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
      cumulative += c;
      progress((int)L + 1, (int)nlaunch, c, cumulative);
    }
  }
  hip_check(hipEventRecord(e1.e, stream_), "hipEventRecord");

  StitchArgs T;
  std::memset(&T, 0, sizeof(T));
  T.summaries = d_sum.as<SegSummary>();
  T.wave_gaps = d_wg.as<uint32_t>();
  T.waves = waves;
  T.nseg = nseg;
  T.seg_bytes = S;
  T.k_base = 0;
  T.seed_last = 5;
  T.seed_max = 2;
  T.large_threshold = large_threshold;
  T.hist = d_hist.as<unsigned long long>();
  T.large = d_large.as<GapRec>();
  T.n_large = d_cnt.as<unsigned int>() + 0;
  T.large_cap = large_cap;
  T.rec = d_rec.as<GapRec>();
  T.n_rec = d_cnt.as<unsigned int>() + 1;
  T.rec_cap = rec_cap;
  T.rescan = d_resc.as<RescanItem>();
  T.n_rescan = d_cnt.as<unsigned int>() + 2;
  T.rescan_cap = rescan_cap;
  T.final_out = d_cnt.as<unsigned long long>() + 4;
  hip_check(launch_stitch(T, stream_), "launch(stitch)");
  hip_check(hipEventRecord(e2.e, stream_), "hipEventRecord");
  {
    // records of flagged wave ranges: the item count stays on the device
    RescanArgs RA;
    RA.items = d_resc.as<RescanItem>();
    RA.n_items = d_cnt.as<unsigned int>() + 2;
    RA.items_cap = rescan_cap;
    RA.out = d_rout.as<GapRec>();
    RA.n_out = d_cnt.as<unsigned int>() + 3;
    RA.out_cap = rout_cap;
    hip_check(launch_rescan(cfg_.block, A, RA, grid_for(rescan_cap), stream_), "launch(rescan)");
  }
  hip_check(hipEventRecord(e3.e, stream_), "hipEventRecord");

  // one gather launch packs counters, totals, the histogram's first
  // kHistHead bins, buckets and the first kHead entries of each list into a
  // staging block, and one copy downloads it (the rest only if needed)
  constexpr uint32_t kHead = 1024;
  constexpr uint32_t kHistHead = 2048;  // gaps below 4096
  const size_t o_cnt = 0, o_fin = 16, o_tot = 32, o_hist = 64;
  const size_t o_bk = o_hist + (size_t)kHistHead * 8;
  const size_t bk_bytes = (size_t)nbuckets * 3 * 8;
  const size_t o_lg = (o_bk + bk_bytes + 255) & ~(size_t)255;
  const size_t o_rc = o_lg + (size_t)kHead * sizeof(GapRec);
  const size_t o_ro = o_rc + (size_t)kHead * sizeof(GapRec);
  const size_t stage_bytes = o_ro + (size_t)kHead * sizeof(GapRec);
  DevBuf& d_stage = ws(kWsStage, stage_bytes);
  {
    SpanList G;
    std::memset(&G, 0, sizeof(G));
    auto gather = [&](size_t off, const void* src, size_t bytes) {
      G.s[G.n++] = CopySpan{static_cast<const uint32_t*>(src), reinterpret_cast<uint32_t*>(d_stage.as<uint8_t>() + off),
                            bytes / 4};
    };
    gather(o_cnt, d_cnt.p, 16);
    gather(o_fin, d_cnt.as<unsigned long long>() + 4, 16);
    gather(o_tot, d_tot.p, 24);
    gather(o_hist, d_hist.p, (size_t)kHistHead * 8);
    if (nbuckets) gather(o_bk, d_bk.p, bk_bytes);
    gather(o_lg, d_large.p, (size_t)kHead * sizeof(GapRec));
    gather(o_rc, d_rec.p, (size_t)kHead * sizeof(GapRec));
    gather(o_ro, d_rout.p, (size_t)kHead * sizeof(GapRec));
    hip_check(launch_spans(G, stream_), "launch(gather)");
  }
  uint8_t* hs = static_cast<uint8_t*>(pinned(stage_bytes));
  hip_check(hipMemcpyAsync(hs, d_stage.p, stage_bytes, hipMemcpyDeviceToHost, stream_), "hipMemcpy");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  const unsigned int* cnts = reinterpret_cast<const unsigned int*>(hs + o_cnt);
  const unsigned long long* fin = reinterpret_cast<const unsigned long long*>(hs + o_fin);
  const unsigned long long* tot = reinterpret_cast<const unsigned long long*>(hs + o_tot);
  const unsigned long long* dh = reinterpret_cast<const unsigned long long*>(hs + o_hist);
  const unsigned long long* bk = reinterpret_cast<const unsigned long long*>(hs + o_bk);
  if (cnts[1] > rec_cap || cnts[2] > rescan_cap || cnts[3] > rout_cap)
    throw Error(7, "internal: record list overflow");  // excluded by the bound above
  const uint32_t n_large = std::min(cnts[0], large_cap), n_rec = std::min(cnts[1], rec_cap),
                 n_rout = std::min(cnts[3], rout_cap);
  auto take = [&](size_t off, const DevBuf& src, uint32_t n) {
    std::vector<GapRec> v(n);
    const uint32_t head = std::min(n, kHead);
    if (head) std::memcpy(v.data(), hs + off, (size_t)head * sizeof(GapRec));
    if (n > head)
      hip_check(hipMemcpy(v.data() + head, src.as<GapRec>() + head, (size_t)(n - head) * sizeof(GapRec),
                          hipMemcpyDeviceToHost), "hipMemcpy");
    return v;
  };
  const std::vector<GapRec> lg = take(o_lg, d_large, n_large), rc = take(o_rc, d_rec, n_rec),
                            ro = take(o_ro, d_rout, n_rout);

  times_ = KernelTimes();
  times_.sieve_ms = elapsed(e0, e1);
  times_.stitch_ms = elapsed(e1, e2);
  times_.rescan_ms = elapsed(e2, e3);
  times_.total_ms = elapsed(e0, e3);
  times_.segments = nseg;

  // ---- assemble (reference conventions)
  const uint64_t cnt = tot[0];
  R.prime_count = nseed + cnt;
  // bins above the largest gap's (fin[1] bounds every gap) are zero; they
  // are downloaded only if that bound reaches past the staged ones
  const bool all_bins = (fin[1] >> 1) >= kHistHead;
  R.hist.assign(2 * (all_bins ? kGapBins : kHistHead), 0);
  for (uint32_t h = 1; h < kHistHead; ++h) R.hist[2 * h] += dh[h];
  if (all_bins) {
    std::vector<unsigned long long> rest(kGapBins - kHistHead);
    hip_check(hipMemcpy(rest.data(), d_hist.as<unsigned long long>() + kHistHead, rest.size() * 8,
                        hipMemcpyDeviceToHost), "hipMemcpy");
    for (uint32_t h = kHistHead; h < kGapBins; ++h) R.hist[2 * h] += rest[h - kHistHead];
  }
  R.hist[1] += 1;  // 2 -> 3
  R.hist[2] += 1;  // 3 -> 5
  R.pi_4_1 = tot[1] + 1;
  R.pi_4_3 = cnt - tot[1] + 1;
  R.pi_3_1 = tot[2];
  R.pi_3_2 = cnt - tot[2] + 2;
  R.last_prime = fin[0];
  for (uint64_t g2 = R.hist.size(); g2-- > 0;)
    if (R.hist[g2]) { R.max_gap = g2; break; }
  R.hist.resize(R.max_gap + 1);
  if (ncp) {
    R.checkpoints.resize(5 * ncp);
    uint64_t c = 0, c41 = 0, c31 = 0;
    for (uint64_t j = 0; j < ncp; ++j) {
      c += bk[3 * j]; c41 += bk[3 * j + 1]; c31 += bk[3 * j + 2];
      R.checkpoints[5 * j + 0] = c + 3;
      R.checkpoints[5 * j + 1] = c41 + 1;
      R.checkpoints[5 * j + 2] = c - c41 + 1;
      R.checkpoints[5 * j + 3] = c31;
      R.checkpoints[5 * j + 4] = c - c31 + 2;
    }
  }
  R.large_total = cnts[0];
  R.large.reserve(lg.size());
  for (const GapRec& g : lg) R.large.push_back({g.position, g.gap});
  std::sort(R.large.begin(), R.large.end(), [](const GapEntry& x, const GapEntry& y) {
    return x.gap != y.gap ? x.gap > y.gap : x.position < y.position;
  });
  std::vector<GapEntry> recs;
  for (const GapRec& g : rc) recs.push_back({g.position, g.gap});
  for (const GapRec& g : ro) recs.push_back({g.position, g.gap});
  std::sort(recs.begin(), recs.end(), [](const GapEntry& x, const GapEntry& y) { return x.position < y.position; });
  for (const GapEntry& g : recs) R.records.push_back(g);
  return R;
}

// ------------------------------------------------------------------ batched ranges
// A batch is one captured hipGraph per capacity class (queries, segments,
// both rounded up to powers of two): H2D of the header, queries and
// segment->query map from pinned staging, a fill kernel for the outputs,
// the persistent batch sieve (dynamic work-list), the per-query stitch, and
// D2H of the results.  Replays only rewrite the pinned staging buffers.
BatchPlan::~BatchPlan() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (h_in) (void)hipHostFree(h_in);
  if (h_out) (void)hipHostFree(h_out);
}

static uint32_t pow2_at_least(uint64_t v, uint32_t lo) {
  uint64_t c = lo;
  while (c < v) c <<= 1;
  if (c > 0x80000000ull) throw Error(kErrArg, "batch too large");
  return (uint32_t)c;
}

std::unique_ptr<BatchPlan> Engine::make_batch_plan(uint32_t qcap, uint32_t scap, const TuneConfig& c,
                                                   hipStream_t st, const uint8_t* cache, uint64_t cache_hi) {
  std::unique_ptr<BatchPlan> up(new BatchPlan());
  BatchPlan& P = *up;
  P.qcap = qcap;
  P.scap = scap;
  P.seg_bytes = c.seg_bytes;
  P.block = c.block;
  P.wgs_per_cu = c.wgs_per_cu;
  P.tier1_p = c.tier1_p;
  P.table_limit = table_limit_;
  P.cache = cache;
  P.cache_hi = cache ? cache_hi : 0;
  P.in_bytes = 64 + (size_t)qcap * sizeof(QueryDesc) + (size_t)scap * 4;
  P.out_bytes = (size_t)qcap * 24 + (size_t)qcap * kQueryBins * 4;
  hip_check(hipHostMalloc(&P.h_in, P.in_bytes, hipHostMallocDefault), "hipHostMalloc");
  hip_check(hipHostMalloc(&P.h_out, P.out_bytes, hipHostMallocDefault), "hipHostMalloc");
  P.d_in = DevBuf(P.in_bytes);
  P.d_sum = DevBuf((size_t)scap * sizeof(SegSummary));
  P.d_out = DevBuf(P.out_bytes);
  P.d_work = DevBuf(256);
  uint8_t* din = P.d_in.as<uint8_t>();
  const uint32_t* d_hdr = reinterpret_cast<const uint32_t*>(din);  // [0] nq, [1] nseg
  const QueryDesc* d_q = reinterpret_cast<const QueryDesc*>(din + 64);
  const uint32_t* d_sq = reinterpret_cast<const uint32_t*>(din + 64 + (size_t)qcap * sizeof(QueryDesc));
  unsigned long long* d_qc = P.d_out.as<unsigned long long>();
  unsigned int* d_qh = reinterpret_cast<unsigned int*>(P.d_out.as<uint8_t>() + (size_t)qcap * 24);

  SieveArgs A = base_args(scap, c);
  A.work = P.d_work.as<unsigned int>();
  A.nseg_dev = d_hdr + 1;
  A.summaries = P.d_sum.as<SegSummary>();
  A.queries = d_q;
  A.seg_query = d_sq;
  A.qhist = d_qh;
  A.qcount = d_qc;
  A.cache = cache;
  BatchStitchArgs B;
  B.queries = d_q;
  B.summaries = P.d_sum.as<SegSummary>();
  B.nq = qcap;
  B.nq_dev = d_hdr;
  B.seg_bytes = c.seg_bytes;
  B.qhist = d_qh;
  B.first = d_qc + qcap;
  B.last = d_qc + 2 * qcap;

  hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  hip_check(hipMemcpyAsync(P.d_in.p, P.h_in, P.in_bytes, hipMemcpyHostToDevice, st), "capture H2D");
  hip_check(launch_zero(P.d_out.as<uint32_t>(), P.out_bytes / 4, st), "capture zero");
  hip_check(launch_zero(P.d_work.as<uint32_t>(), 64, st), "capture zero");
  hip_check(launch_sieve(kModeBatch, c.block, A, grid_for(scap, c), st), "capture sieve");
  hip_check(launch_batch_stitch(B, qcap, st), "capture stitch");
  hip_check(hipMemcpyAsync(P.h_out, P.d_out.p, P.out_bytes, hipMemcpyDeviceToHost, st), "capture D2H");
  hip_check(hipStreamEndCapture(st, &P.graph), "hipStreamEndCapture");
  hip_check(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  return up;
}

BatchPlan& Engine::batch_plan(uint32_t nq, uint32_t nseg) {
  const uint32_t qcap = pow2_at_least(nq, 64), scap = pow2_at_least(nseg, 1024);
  for (auto& p : plans_)
    if (p->qcap == qcap && p->scap == scap && p->seg_bytes == cfg_.seg_bytes && p->block == cfg_.block &&
        p->wgs_per_cu == cfg_.wgs_per_cu && p->tier1_p == cfg_.tier1_p && p->table_limit == table_limit_)
      return *p;
  if (plans_.size() >= 8) plans_.erase(plans_.begin());
  plans_.push_back(make_batch_plan(qcap, scap, cfg_, stream_));
  return *plans_.back();
}

Engine::BatchLayout Engine::layout_batch(const uint64_t* a, const uint64_t* b, size_t n, uint32_t S) const {
  BatchLayout L;
  for (size_t i = 0; i < n; ++i) {
    const uint64_t lo = std::max<uint64_t>(a[i], 7);
    if (b[i] < lo) continue;
    L.order.push_back((uint32_t)i);
  }
  // per-segment cost grows with sqrt(b): deal the expensive queries first
  std::stable_sort(L.order.begin(), L.order.end(), [&](uint32_t x, uint32_t y) { return b[x] > b[y]; });
  L.nq = (uint32_t)L.order.size();
  uint64_t total = 0;
  for (uint32_t q = 0; q < L.nq; ++q) {
    const uint64_t lo = std::max<uint64_t>(a[L.order[q]], 7);
    total += ((b[L.order[q]] / 30 + 1) - lo / 30 + S - 1) / S;
  }
  if (total > 0x40000000ull) throw Error(kErrArg, "batch too large");
  L.nseg = (uint32_t)total;
  return L;
}

void Engine::stage_batch(BatchPlan& P, const BatchLayout& L, const uint64_t* a, const uint64_t* b) {
  if (L.nq > P.qcap || L.nseg > P.scap) throw Error(kErrInternal, "batch exceeds its plan");
  const uint32_t S = P.seg_bytes;
  uint8_t* hin = static_cast<uint8_t*>(P.h_in);
  uint32_t* hdr = reinterpret_cast<uint32_t*>(hin);
  QueryDesc* hq = reinterpret_cast<QueryDesc*>(hin + 64);
  uint32_t* hsq = reinterpret_cast<uint32_t*>(hin + 64 + (size_t)P.qcap * sizeof(QueryDesc));
  hdr[0] = L.nq;
  hdr[1] = L.nseg;
  uint32_t seg = 0;
  for (uint32_t q = 0; q < L.nq; ++q) {
    const uint32_t i = L.order[q];
    QueryDesc& Q = hq[q];
    Q.a = std::max<uint64_t>(a[i], 7);
    Q.b = b[i];
    Q.k_lo = Q.a / 30;
    Q.k_hi = b[i] / 30 + 1;
    Q.seg_first = seg;
    Q.nseg = (uint32_t)((Q.k_hi - Q.k_lo + S - 1) / S);
    Q.from_cache = (P.cache && Q.b <= P.cache_hi) ? 1u : 0u;
    Q.pad = 0;
    for (uint32_t k = 0; k < Q.nseg; ++k) hsq[seg + k] = q;
    seg += Q.nseg;
  }
}

void Engine::read_batch(const BatchPlan* P, const BatchLayout& L, const uint64_t* a, const uint64_t* b, size_t n,
                        RangeOut* out) {
  // every record written once (out may be uninitialised): zeros for the
  // queries with no wheel bytes (those layout_batch leaves out), the device
  // results for the others
  for (size_t i = 0; i < n; ++i)
    if (b[i] < std::max<uint64_t>(a[i], 7)) out[i] = RangeOut();
  if (L.nq) {
    const unsigned long long* qc = static_cast<const unsigned long long*>(P->h_out);
    const uint32_t* qh = reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(P->h_out) + (size_t)P->qcap * 24);
    for (uint32_t q = 0; q < L.nq; ++q) {
      RangeOut& o = out[L.order[q]];
      o.count = qc[q];
      o.first = qc[P->qcap + q];
      o.last = qc[2 * P->qcap + q];
      std::memcpy(o.hist, qh + (size_t)q * kQueryBins, kQueryBins * 4);
      o.pad = 0;
    }
  }
  // seeds 2, 3, 5 (not in the wheel bitmap) for queries that include them
  for (size_t i = 0; i < n; ++i) {
    if (a[i] > 5 || b[i] < 2 || b[i] < a[i]) continue;
    RangeOut& o = out[i];
    uint64_t sd[3];
    int ns = 0;
    for (uint64_t s : {2ull, 3ull, 5ull}) if (s >= a[i] && s <= b[i]) sd[ns++] = s;
    if (!ns) continue;
    auto add_gap = [&](uint64_t g) {
      if (g >= 2 && g <= 512 && !(g & 1)) o.hist[g / 2]++; else o.hist[0]++;
    };
    for (int k = 1; k < ns; ++k) add_gap(sd[k] - sd[k - 1]);
    if (o.count) add_gap(o.first - sd[ns - 1]);
    else o.last = sd[ns - 1];
    o.first = sd[0];
    o.count += ns;
  }
}

void Engine::batch_ranges(const uint64_t* a, const uint64_t* b, size_t n, RangeOut* out) {
  cfg_ = cfgs_[1];
  BatchLayout L = layout_batch(a, b, n, cfg_.seg_bytes);
  times_ = KernelTimes();
  if (L.nq) {
    uint64_t bmax = 7;
    for (uint32_t i : L.order) bmax = std::max(bmax, b[i]);
    ensure_tables(bmax);
    BatchPlan& P = batch_plan(L.nq, L.nseg);
    stage_batch(P, L, a, b);
    Event e0, e1;
    hip_check(hipEventRecord(e0.e, stream_), "hipEventRecord");
    hip_check(hipGraphLaunch(P.exec, stream_), "hipGraphLaunch");
    hip_check(hipEventRecord(e1.e, stream_), "hipEventRecord");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    times_.sieve_ms = elapsed(e0, e1);
    times_.total_ms = times_.sieve_ms;
    times_.segments = L.nseg;
    read_batch(&P, L, a, b, n, out);
  } else {
    read_batch(nullptr, L, a, b, n, out);
  }
}

// ------------------------------------------------------------------ bitmap + list
uint64_t Engine::bitmap_range(uint64_t lo, uint64_t hi, DevBuf& out, uint64_t* k_base_out) {
  cfg_ = cfgs_[0];
  if (lo < 7) lo = 7;
  const uint32_t S = cfg_.seg_bytes;
  const uint64_t k_base = lo / 30, k_end = hi / 30 + 1;
  *k_base_out = k_base;
  if (hi < lo) return 0;
  ensure_tables(hi);
  const uint64_t nbytes = k_end - k_base;
  const uint64_t nseg64 = (nbytes + S - 1) / S;
  if (nseg64 > 0x7FFFFFF0ull) throw Error(kErrArg, "range too large");
  const uint32_t nseg = (uint32_t)nseg64;
  if (out.bytes < nseg64 * S + 64) out = DevBuf(nseg64 * S + 64);
  DevBuf& wk = ws(kWsWork, 64);
  hip_check(hipMemsetAsync(wk.p, 0, 64, stream_), "memset");
  SieveArgs A = base_args(nseg);
  A.k_base = k_base;
  A.k_end = k_end;
  A.lo_val = lo;
  A.hi_val = hi;
  A.work = wk.as<unsigned int>();
  A.bitmap = out.as<uint8_t>();
  Event e0, e1;
  hip_check(hipEventRecord(e0.e, stream_), "hipEventRecord");
  hip_check(launch_sieve(kModeBitmap, cfg_.block, A, grid_for(nseg), stream_), "launch(bitmap)");
  hip_check(hipEventRecord(e1.e, stream_), "hipEventRecord");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  times_ = KernelTimes();
  times_.sieve_ms = elapsed(e0, e1);
  times_.total_ms = times_.sieve_ms;
  times_.segments = nseg;
  return nbytes;
}

std::vector<uint64_t> Engine::primes_in_range(uint64_t a, uint64_t b, uint64_t max_count, bool* overflow) {
  std::vector<uint64_t> res;
  *overflow = false;
  for (uint64_t sd : {2ull, 3ull, 5ull})
    if (sd >= a && sd <= b) res.push_back(sd);
  const uint64_t lo = std::max<uint64_t>(a, 7);
  if (b < lo) {
    if (res.size() > max_count) *overflow = true;
    return res;
  }
  DevBuf bm;
  uint64_t k_base = 0;
  const uint64_t nbytes = bitmap_range(lo, b, bm, &k_base);
  const uint64_t nblk = (nbytes + kListBlockBytes - 1) / kListBlockBytes;
  DevBuf d_cnt(nblk * 4 + 4), d_off(nblk * 8 + 8);
  hip_check(launch_list_count(bm.as<uint8_t>(), nbytes, d_cnt.as<uint32_t>(), stream_), "launch(list_count)");
  std::vector<uint32_t> cnt(nblk);
  hip_check(hipMemcpyAsync(cnt.data(), d_cnt.p, nblk * 4, hipMemcpyDeviceToHost, stream_), "D2H");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  std::vector<unsigned long long> off(nblk);
  uint64_t total = 0;
  for (uint64_t i = 0; i < nblk; ++i) { off[i] = total; total += cnt[i]; }
  if (res.size() + total > max_count) {
    *overflow = true;
    return res;
  }
  if (total) {
    DevBuf d_out(total * 8);
    hip_check(hipMemcpyAsync(d_off.p, off.data(), nblk * 8, hipMemcpyHostToDevice, stream_), "H2D");
    hip_check(launch_list_write(bm.as<uint8_t>(), nbytes, k_base, d_off.as<unsigned long long>(),
                                d_out.as<uint64_t>(), stream_), "launch(list_write)");
    const size_t s0 = res.size();
    res.resize(s0 + total);
    hip_check(hipMemcpyAsync(res.data() + s0, d_out.p, total * 8, hipMemcpyDeviceToHost, stream_), "D2H");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }
  return res;
}

}  // namespace voss
