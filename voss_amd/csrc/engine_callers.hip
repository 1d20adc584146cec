// engine_callers.hip — the callers around the sieve: Goldbach, Sophie Germain,
// nth prime and batched Miller-Rabin, all on the device.
//
// Reference entry points restated here (library/src/primes/):
//   goldbach_upto.cu:48-150   (bitmap + one thread per even n)
//   goldbach.cpp:21-69        (partition counts; here from the bitmap)
//   context.cu:765-888 + sophie_germain_upto.cu:22-33
//   nth.cu:40-103             (here: per-segment counts, then one segment's list)
//   next_prev.cu, core/miller_rabin.cpp (here: batched device MR)
#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.h"
#include "kernels.h"
#include "misc_kernels.h"

namespace voss {

namespace {
constexpr int kErrArg = 6, kErrInternal = 7;
}

// Even n in [6, N] with no partition p1 + p2 = n, p1 <= p2, p1 <= pmax (0:
// no limit; a limit plants counterexamples to test with), ascending, the
// first max_ce of them (every one counted into *total when given; n = 4 =
// 2 + 2 always has one).  A first pass (goldbach_check_kernel) tries the
// odd primes up to min(N/2, pmax, scan_limit if set, the sieving table's
// 2^24); an n left without a partner is final when that list held every
// partner <= min(n/2, pmax), and is otherwise settled by an exact count of
// its partitions over those partners.  Candidates arrive in windows of at
// most kCap even n when the whole range yields more than kCap of them.
std::vector<uint64_t> Engine::goldbach_upto(uint64_t N, uint64_t max_ce, uint64_t pmax, uint64_t scan_limit,
                                            uint64_t* total) {
  std::vector<uint64_t> ce;
  if (total) *total = 0;
  if (N < 6 || (max_ce == 0 && !total)) return ce;
  const uint64_t n_top = N & ~1ull;
  const uint64_t plim = pmax ? pmax : ~0ull;
  DevBuf bm;
  uint64_t k_base = 0;
  bitmap_range(7, N, bm, &k_base);  // k_base = 0
  uint64_t lim = std::min<uint64_t>(n_top / 2, plim);
  if (scan_limit) lim = std::min(lim, scan_limit);
  lim = std::min<uint64_t>(lim, table_limit_);
  // the odd primes below the sieving table's first prime, then the table
  std::vector<uint32_t> sp;
  for (uint32_t q = 3; q < kFirstSievingPrime && q <= lim; q += 2) {
    bool pr = true;
    for (uint32_t d = 3; d * d <= q; d += 2) pr = pr && (q % d != 0);
    if (pr) sp.push_back(q);
  }
  for (uint32_t p : h_primes_) {
    if ((uint64_t)p > lim) break;
    sp.push_back(p);
  }
  constexpr uint64_t kCap = 1u << 20;
  DevBuf d_sp(sp.size() * 4 + 4), d_ce(kCap * 8), d_n(8);
  if (!sp.empty())
    hip_check(hipMemcpyAsync(d_sp.p, sp.data(), sp.size() * 4, hipMemcpyHostToDevice, stream_), "H2D");
  auto pass = [&](uint64_t lo, uint64_t hi) -> uint64_t {
    hip_check(hipMemsetAsync(d_n.p, 0, 8, stream_), "memset");
    hip_check(launch_goldbach_check(bm.as<uint8_t>(), d_sp.as<uint32_t>(), (uint32_t)sp.size(), lo, hi,
                                    d_ce.as<uint64_t>(), d_n.as<unsigned long long>(), kCap, stream_),
              "launch(goldbach_check)");
    unsigned long long n = 0;
    hip_check(hipMemcpyAsync(&n, d_n.p, 8, hipMemcpyDeviceToHost, stream_), "D2H");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    return n;
  };
  // take the candidates of one pass in order; true when the caller needs no more
  auto take = [&](uint64_t n) -> bool {
    if (!n) return false;
    std::vector<uint64_t> cand(n);
    hip_check(hipMemcpy(cand.data(), d_ce.p, n * 8, hipMemcpyDeviceToHost), "D2H");
    std::sort(cand.begin(), cand.end());
    // n is final when the scan held every partner <= min(n/2, pmax)
    std::vector<uint64_t> open;
    for (uint64_t c : cand)
      if (std::min(c / 2, plim) > lim) open.push_back(c);
    std::vector<uint64_t> cnt;
    if (!open.empty()) cnt = goldbach_counts_bm(bm, open.data(), open.size(), N, plim);
    size_t j = 0;
    for (uint64_t c : cand) {
      bool none = true;
      if (j < open.size() && open[j] == c) none = cnt[j++] == 0;
      if (!none) continue;
      if (total) ++*total;
      if (ce.size() < max_ce) ce.push_back(c);
      else if (!total) return true;
    }
    return !total && ce.size() >= max_ce;
  };
  const uint64_t all = pass(6, n_top);
  if (all <= kCap) {
    take(all);
    return ce;
  }
  for (uint64_t lo = 6; lo <= n_top; lo += 2 * kCap) {
    const uint64_t hi = std::min<uint64_t>(n_top, lo + 2 * (kCap - 1));
    if (take(pass(lo, hi))) break;
  }
  return ce;
}

std::vector<uint64_t> Engine::goldbach_counts_bm(const DevBuf& bm, const uint64_t* ns, size_t k, uint64_t nmax,
                                                 uint64_t plim) {
  std::vector<uint64_t> out(k, 0);
  if (!k) return out;
  DevBuf d_ns(k * 8), d_out(k * 8);
  hip_check(hipMemcpyAsync(d_ns.p, ns, k * 8, hipMemcpyHostToDevice, stream_), "H2D");
  for (size_t i0 = 0; i0 < k; i0 += 65535)
    hip_check(launch_goldbach_count(bm.as<uint8_t>(), d_ns.as<uint64_t>() + i0, std::min<size_t>(65535, k - i0), nmax,
                                    plim, d_out.as<unsigned long long>() + i0, stream_),
              "launch(goldbach_count)");
  hip_check(hipMemcpyAsync(out.data(), d_out.p, k * 8, hipMemcpyDeviceToHost, stream_), "D2H");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  return out;
}

std::vector<uint64_t> Engine::goldbach_counts(const uint64_t* ns, size_t k) {
  std::vector<uint64_t> out(k, 0);
  if (!k) return out;
  uint64_t nmax = 7;
  for (size_t i = 0; i < k; ++i) {
    if (ns[i] < 4 || (ns[i] & 1)) throw Error(1, "n must be even and >= 4");
    nmax = std::max(nmax, ns[i]);
  }
  DevBuf bm;
  uint64_t k_base = 0;
  bitmap_range(7, nmax, bm, &k_base);
  return goldbach_counts_bm(bm, ns, k, nmax, ~0ull);
}

uint64_t Engine::sophie_germain(uint64_t N) {
  if (N < 2) return 0;
  uint64_t c = 1;            // 2 -> 5
  if (N >= 3) ++c;           // 3 -> 7
  if (N >= 5) ++c;           // 5 -> 11
  if (N < 7) return c;
  DevBuf bm;
  uint64_t k_base = 0;
  bitmap_range(7, 2 * N + 1, bm, &k_base);
  DevBuf d_c(8);
  hip_check(hipMemsetAsync(d_c.p, 0, 8, stream_), "memset");
  hip_check(launch_sophie(bm.as<uint8_t>(), N, d_c.as<unsigned long long>(), stream_), "launch(sophie)");
  unsigned long long s = 0;
  hip_check(hipMemcpyAsync(&s, d_c.p, 8, hipMemcpyDeviceToHost, stream_), "D2H");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  return c + s;
}

void Engine::is_prime_batch(const uint64_t* x, size_t n, uint8_t* out) {
  if (!n) return;
  DevBuf d_x(n * 8), d_o(n + 8);
  hip_check(hipMemcpyAsync(d_x.p, x, n * 8, hipMemcpyHostToDevice, stream_), "H2D");
  hip_check(launch_is_prime(d_x.as<uint64_t>(), n, d_o.as<uint8_t>(), stream_), "launch(is_prime)");
  hip_check(hipMemcpyAsync(out, d_o.p, n, hipMemcpyDeviceToHost, stream_), "D2H");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Engine::next_prev_batch(const uint64_t* x, size_t n, uint64_t* out, int dir) {
  if (!n) return;
  DevBuf d_x(n * 8), d_o(n * 8);
  hip_check(hipMemcpyAsync(d_x.p, x, n * 8, hipMemcpyHostToDevice, stream_), "H2D");
  hip_check(launch_next_prev(d_x.as<uint64_t>(), n, d_o.as<uint64_t>(), dir, stream_), "launch(next_prev)");
  hip_check(hipMemcpyAsync(out, d_o.p, n * 8, hipMemcpyDeviceToHost, stream_), "D2H");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

uint32_t Engine::segment_counts(uint64_t hi, std::vector<uint32_t>& counts, const TuneConfig& c) {
  cfg_ = c;
  ensure_tables(hi);
  const uint32_t S = cfg_.seg_bytes;
  const uint64_t k_end = hi / 30 + 1;
  const uint64_t nseg64 = (k_end + S - 1) / S;
  if (nseg64 > 0x7FFFFFF0ull) throw Error(kErrArg, "range too large");
  const uint32_t nseg = (uint32_t)nseg64;
  DevBuf& wk = ws(kWsWork, 64);
  DevBuf& tot = ws(kWsTotals, 64);
  DevBuf& sum = ws(kWsSummaries, (size_t)nseg * sizeof(SegSummary));
  hip_check(hipMemsetAsync(wk.p, 0, 64, stream_), "memset");
  hip_check(hipMemsetAsync(tot.p, 0, 64, stream_), "memset");
  SieveArgs A = base_args(nseg);
  A.k_base = 0;
  A.k_end = k_end;
  A.lo_val = 7;
  A.hi_val = hi;
  A.work = wk.as<unsigned int>();
  A.totals = tot.as<unsigned long long>();
  A.summaries = sum.as<SegSummary>();
  hip_check(launch_sieve(kModeCount, cfg_.block, A, grid_for(nseg), stream_), "launch(count)");
  std::vector<SegSummary> ss(nseg);
  hip_check(hipMemcpyAsync(ss.data(), sum.p, (size_t)nseg * sizeof(SegSummary), hipMemcpyDeviceToHost, stream_),
            "D2H");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  counts.resize(nseg);
  for (uint32_t i = 0; i < nseg; ++i) counts[i] = ss[i].count;
  return nseg;
}

std::shared_ptr<const NthIndex> Engine::nth_index(uint64_t hi, const TuneConfig& c) {
  if (nth_cache_ && nth_cache_->hi >= hi && nth_cache_->seg_bytes == c.seg_bytes) return nth_cache_;
  std::shared_ptr<NthIndex> idx(new NthIndex());
  std::vector<uint32_t> counts;
  const uint32_t nseg = segment_counts(hi, counts, c);
  idx->hi = hi;
  idx->seg_bytes = c.seg_bytes;
  idx->cum.resize((size_t)nseg + 1);
  uint64_t t = 3;  // 2, 3, 5
  for (uint32_t s = 0; s < nseg; ++s) {
    idx->cum[s] = t;
    t += counts[s];
  }
  idx->cum[nseg] = t;
  nth_cache_ = idx;
  return nth_cache_;
}

uint64_t Engine::nth_prime(uint64_t n) {
  if (n == 0) throw Error(1, "n must be >= 1 (1-indexed)");
  static const uint64_t small[6] = {2, 3, 5, 7, 11, 13};
  if (n <= 6) return small[n - 1];
  const double nf = (double)n;
  // Rosser: p_n < n (ln n + ln ln n) for n >= 6
  const uint64_t hi = (uint64_t)(nf * (std::log(nf) + std::log(std::log(nf)))) + 64;
  uint64_t cum = 3;  // 2, 3, 5
  uint32_t s = 0;
  if (nth_cache_ && nth_cache_->hi >= hi && nth_cache_->seg_bytes == cfgs_[0].seg_bytes) {
    // a resident index (built by a request server) already covers p_n
    const std::vector<uint64_t>& c = nth_cache_->cum;
    s = (uint32_t)(std::lower_bound(c.begin(), c.end(), n) - c.begin()) - 1;
    cum = c[s];
  } else {
    std::vector<uint32_t> counts;
    const uint32_t nseg = segment_counts(hi, counts, cfgs_[0]);
    for (; s < nseg; ++s) {
      if (cum + counts[s] >= n) break;
      cum += counts[s];
    }
    if (s == nseg) throw Error(kErrInternal, "internal: could not locate nth prime");
  }
  const uint64_t S = cfgs_[0].seg_bytes;
  const uint64_t lo = std::max<uint64_t>(30 * (uint64_t)s * S, 7);
  const uint64_t top = std::min<uint64_t>(30 * ((uint64_t)s + 1) * S - 1, hi);
  bool of = false;
  std::vector<uint64_t> seg = primes_in_range(lo, top, ~0ull, &of);
  const uint64_t need = n - cum;
  if (need == 0 || need > seg.size()) throw Error(kErrInternal, "internal: nth prime segment mismatch");
  return seg[need - 1];
}

}  // namespace voss
