// capi.cpp — the voss C ABI over the MI355X engine.
//
// Signatures, codes, argument checks (and their order), ownership and
// caching follow the reference entry points cited per function (paths under
// library/src/).  No exception crosses the ABI: voss::Error carries its
// VOSS_* code, std::bad_alloc maps to VOSS_ERR_OUT_OF_MEMORY and any other
// std::exception to VOSS_ERR_CUDA, as context.cu:628-634 does.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "engine.h"
#include "host_math.h"
#include "serve.h"
#include "voss/voss.h"
#include "voss/voss_amd.h"
#include "voss/voss_numbers.h"
#include "voss/voss_primes.h"

using voss::Engine;
using voss::Error;

namespace {

thread_local std::string g_last_error;
void set_err(const std::string& m) { g_last_error = m; }

constexpr uint64_t kMaxN = 100000000000000ull;  // 1e14 (context.cu:578, prime_count.cu:303)
constexpr uint32_t kLargeThreshold = 500;        // context.cu:64
constexpr int kMaxGap = 100000;                  // context.cu:63

template <typename F>
int guarded(F&& f, int generic = VOSS_ERR_CUDA) {
  try {
    return f();
  } catch (const Error& e) {
    set_err(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_err("host allocation failed");
    return VOSS_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    set_err(e.what());
    return generic;
  } catch (...) {
    set_err("Unknown error");
    return VOSS_ERR_INTERNAL;
  }
}

Engine& eng() { return Engine::get(); }

// malloc'd copy for arrays handed to the caller (freed by voss_free).
uint64_t* to_malloc(const std::vector<uint64_t>& v) {
  uint64_t* p = static_cast<uint64_t*>(std::malloc(v.size() * sizeof(uint64_t)));
  if (!p) throw Error(VOSS_ERR_OUT_OF_MEMORY, "host malloc failed");
  std::memcpy(p, v.data(), v.size() * sizeof(uint64_t));
  return p;
}

}  // namespace

namespace voss {
void numbers_set_err(const char* m) { set_err(m); }
}

// ====================================================================== errors
extern "C" const char* voss_get_last_error(void) { return g_last_error.c_str(); }

extern "C" const char* voss_strerror(int code) {  // error.cpp:20-32
  switch (code) {
    case VOSS_OK: return "OK";
    case VOSS_ERR_INVALID_N: return "Invalid N";
    case VOSS_ERR_OUT_OF_RANGE: return "Out of range";
    case VOSS_ERR_NO_CUDA: return "No CUDA device";
    case VOSS_ERR_OUT_OF_MEMORY: return "Out of memory";
    case VOSS_ERR_CUDA: return "CUDA error";
    case VOSS_ERR_INVALID_ARG: return "Invalid argument";
    case VOSS_ERR_INTERNAL: return "Internal error";
    default: return "Unknown error";
  }
}

extern "C" void voss_free(void* ptr) { std::free(ptr); }

// ====================================================================== prime_count
extern "C" int voss_primes_prime_count(uint64_t N, uint64_t* out) {  // prime_count.cu:298-322
  if (N < 2) { set_err("N must be >= 2"); return VOSS_ERR_INVALID_N; }
  if (N > kMaxN) { set_err("N exceeds M0 maximum (10^14)"); return VOSS_ERR_OUT_OF_RANGE; }
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (N < 7) { *out = N < 3 ? 1 : (N < 5 ? 2 : 3); return VOSS_OK; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    *out = e.prime_count(N);
    return VOSS_OK;
  });
}

// ====================================================================== context
struct voss_primes_ctx {
  uint64_t N = 0;
  int profile = VOSS_PROFILE_STANDARD;
  bool has_count = false;
  uint64_t count = 0;
  bool has_full = false;
  std::vector<int64_t> hist;
  uint64_t c41 = 0, c43 = 0;
  std::vector<voss::GapEntry> large;
  bool has_sophie = false;
  uint64_t sophie = 0;
  voss_progress_cb cb = nullptr;
  void* user = nullptr;
};

extern "C" int voss_primes_ctx_new(uint64_t N, int profile, voss_primes_ctx** out_ctx) {  // context.cu:566-599
  if (!out_ctx) { set_err("out_ctx pointer is null"); return VOSS_ERR_INVALID_ARG; }
  *out_ctx = nullptr;
  if (N < 2) { set_err("N must be >= 2"); return VOSS_ERR_INVALID_N; }
  if (N > kMaxN) { set_err("N exceeds maximum (10^14)"); return VOSS_ERR_OUT_OF_RANGE; }
  if (profile != VOSS_PROFILE_MINIMAL && profile != VOSS_PROFILE_STANDARD && profile != VOSS_PROFILE_FULL) {
    set_err("invalid profile value");
    return VOSS_ERR_INVALID_ARG;
  }
  return guarded([&] {
    voss_primes_ctx* c = new voss_primes_ctx();
    c->N = N;
    c->profile = profile;
    *out_ctx = c;
    return VOSS_OK;
  }, VOSS_ERR_INTERNAL);
}

extern "C" void voss_primes_ctx_free(voss_primes_ctx* ctx) { delete ctx; }

// Lazy computation (context.cu:608-635): MINIMAL caches pi only; STANDARD and
// FULL compute the whole statistics set in one sweep.
static int ensure_computed(voss_primes_ctx* ctx, bool need_full) {
  if (ctx->has_count && (!need_full || ctx->has_full)) return VOSS_OK;
  const bool full = ctx->profile != VOSS_PROFILE_MINIMAL;
  if (ctx->N < 7) {  // context.cu:228-247 needs no device
    voss::SweepStats S = voss::small_sweep_stats(ctx->N);
    ctx->count = S.prime_count;
    ctx->has_count = true;
    ctx->hist.assign(kMaxGap, 0);
    for (size_t g = 0; g < std::min<size_t>(kMaxGap, S.hist.size()); ++g) ctx->hist[g] = (int64_t)S.hist[g];
    ctx->c41 = ctx->c43 = 0;
    ctx->has_full = full;
    return VOSS_OK;
  }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    voss::ProgressFn pf;
    if (ctx->cb) {
      voss_progress_cb cb = ctx->cb;
      void* user = ctx->user;
      pf = [cb, user](int seg, int total, uint64_t in_seg, uint64_t cum) { cb(seg, total, in_seg, cum, user); };
    }
    if (!full && !pf) {
      ctx->count = e.prime_count(ctx->N);
      ctx->has_count = true;
      return VOSS_OK;
    }
    voss::SweepStats S = e.sweep_stats(ctx->N, 0, kLargeThreshold, pf);
    ctx->count = S.prime_count;
    ctx->has_count = true;
    if (full) {
      ctx->hist.assign(kMaxGap, 0);
      for (size_t g = 0; g < std::min<size_t>(kMaxGap, S.hist.size()); ++g) ctx->hist[g] = (int64_t)S.hist[g];
      ctx->c41 = S.pi_4_1;
      ctx->c43 = S.pi_4_3;
      ctx->large = std::move(S.large);
      ctx->has_full = true;
    }
    return VOSS_OK;
  });
}

extern "C" int voss_primes_ctx_set_progress(voss_primes_ctx* ctx, voss_progress_cb cb, void* user) {
  if (!ctx) { set_err("ctx is null"); return VOSS_ERR_INVALID_ARG; }
  ctx->cb = cb;
  ctx->user = user;
  return VOSS_OK;
}

extern "C" int voss_primes_ctx_prime_count(voss_primes_ctx* ctx, uint64_t* out) {  // context.cu:652-667
  if (!ctx) { set_err("ctx is null"); return VOSS_ERR_INVALID_ARG; }
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (!ctx->has_count) {
    int rc = ensure_computed(ctx, ctx->profile != VOSS_PROFILE_MINIMAL);
    if (rc != VOSS_OK) return rc;
  }
  *out = ctx->count;
  return VOSS_OK;
}

static int require_full(voss_primes_ctx* ctx, const char* what) {
  if (ctx->profile == VOSS_PROFILE_MINIMAL) {
    set_err(std::string("Profile MINIMAL does not compute ") + what + "; use STANDARD or FULL");
    return VOSS_ERR_INVALID_ARG;
  }
  if (!ctx->has_full) return ensure_computed(ctx, true);
  return VOSS_OK;
}

static int gap_count(voss_primes_ctx* ctx, uint64_t* out, int g) {  // context.cu:669-689
  if (!ctx) { set_err("ctx is null"); return VOSS_ERR_INVALID_ARG; }
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  int rc = require_full(ctx, "gap histogram");
  if (rc != VOSS_OK) return rc;
  *out = (uint64_t)ctx->hist[g];
  return VOSS_OK;
}
extern "C" int voss_primes_ctx_twins(voss_primes_ctx* ctx, uint64_t* out) { return gap_count(ctx, out, 2); }
extern "C" int voss_primes_ctx_cousin(voss_primes_ctx* ctx, uint64_t* out) { return gap_count(ctx, out, 4); }
extern "C" int voss_primes_ctx_sexy(voss_primes_ctx* ctx, uint64_t* out) { return gap_count(ctx, out, 6); }

extern "C" int voss_primes_ctx_statistics(voss_primes_ctx* ctx, voss_primes_stats_t* out) {  // context.cu:694-755
  if (!ctx) { set_err("ctx is null"); return VOSS_ERR_INVALID_ARG; }
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  int rc = require_full(ctx, "gap histogram");
  if (rc != VOSS_OK) return rc;
  const std::vector<int64_t>& h = ctx->hist;
  // same accumulation order as the reference, so the doubles match bit for bit
  int64_t total = 0, weighted = 0;
  for (int g = 1; g < kMaxGap; ++g) {
    if (h[g] == 0) continue;
    total += h[g];
    weighted += (int64_t)g * h[g];
  }
  if (total == 0) {
    out->mean_gap = out->std_dev = out->skewness = out->kurtosis = 0.0;
    out->total_gaps = 0;
    return VOSS_OK;
  }
  const double mean = (double)weighted / (double)total;
  double var_sum = 0.0, m3 = 0.0, m4 = 0.0;
  for (int g = 1; g < kMaxGap; ++g) {
    if (h[g] == 0) continue;
    const double d = (double)g - mean;
    const double d2 = d * d;
    var_sum += (double)h[g] * d2;
    m3 += (double)h[g] * d2 * d;
    m4 += (double)h[g] * d2 * d2;
  }
  const double var = var_sum / (double)total;
  const double sd = std::sqrt(var);
  out->mean_gap = mean;
  out->std_dev = sd;
  out->skewness = (sd > 0.0) ? (m3 / (double)total) / (var * sd) : 0.0;
  out->kurtosis = (var > 0.0) ? (m4 / (double)total) / (var * var) - 3.0 : 0.0;
  out->total_gaps = (uint64_t)total;
  return VOSS_OK;
}

extern "C" int voss_primes_ctx_chebyshev_bias(voss_primes_ctx* ctx, uint64_t* out_pi_4_1, uint64_t* out_pi_4_3) {
  if (!ctx) { set_err("ctx is null"); return VOSS_ERR_INVALID_ARG; }  // context.cu:940-963
  if (!out_pi_4_1 || !out_pi_4_3) { set_err("out pointers must not be null"); return VOSS_ERR_INVALID_ARG; }
  int rc = require_full(ctx, "Chebyshev bias");
  if (rc != VOSS_OK) return rc;
  *out_pi_4_1 = ctx->c41;
  *out_pi_4_3 = ctx->c43;
  return VOSS_OK;
}

extern "C" int voss_primes_ctx_large_gaps_count(voss_primes_ctx* ctx, uint64_t* out_count) {  // context.cu:896-912
  if (!ctx || !out_count) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  if (ctx->profile == VOSS_PROFILE_MINIMAL) { set_err("Profile MINIMAL does not compute large gaps"); return VOSS_ERR_INVALID_ARG; }
  int rc = require_full(ctx, "large gaps");
  if (rc != VOSS_OK) return rc;
  *out_count = ctx->large.size();
  return VOSS_OK;
}

extern "C" int voss_primes_ctx_large_gaps_get(voss_primes_ctx* ctx, uint64_t index, uint64_t* out_position,
                                              uint32_t* out_gap) {  // context.cu:914-938
  if (!ctx || !out_position || !out_gap) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  if (ctx->profile == VOSS_PROFILE_MINIMAL) { set_err("Profile MINIMAL does not compute large gaps"); return VOSS_ERR_INVALID_ARG; }
  int rc = require_full(ctx, "large gaps");
  if (rc != VOSS_OK) return rc;
  if (index >= ctx->large.size()) { set_err("index out of range"); return VOSS_ERR_INVALID_ARG; }
  *out_position = ctx->large[index].position;
  *out_gap = ctx->large[index].gap;
  return VOSS_OK;
}

extern "C" int voss_primes_ctx_sophie_germain_upto(voss_primes_ctx* ctx, uint64_t* out) {  // context.cu:765-888
  if (!ctx) { set_err("ctx is null"); return VOSS_ERR_INVALID_ARG; }
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (ctx->profile == VOSS_PROFILE_MINIMAL) {
    set_err("Profile MINIMAL does not support sophie_germain; use STANDARD or FULL");
    return VOSS_ERR_INVALID_ARG;
  }
  if (ctx->has_sophie) { *out = ctx->sophie; return VOSS_OK; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    ctx->sophie = e.sophie_germain(ctx->N);
    ctx->has_sophie = true;
    *out = ctx->sophie;
    return VOSS_OK;
  }, VOSS_ERR_INTERNAL);
}

extern "C" int voss_primes_ctx_sophie_germain(voss_primes_ctx* ctx, uint64_t* out) {
  return voss_primes_ctx_sophie_germain_upto(ctx, out);  // context.cu:890-894
}

extern "C" int voss_primes_ctx_export_csv(voss_primes_ctx* ctx, const char* prefix) {  // io/csv.cpp:17-115
  if (!ctx) { set_err("ctx is null"); return VOSS_ERR_INVALID_ARG; }
  if (!prefix) { set_err("prefix is null"); return VOSS_ERR_INVALID_ARG; }
  const std::string p(prefix);
  uint64_t c1 = 0, c3 = 0;
  int rc = voss_primes_ctx_chebyshev_bias(ctx, &c1, &c3);
  if (rc != VOSS_OK) return rc;
  {
    const std::string path = p + "_chebyshev.csv";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) { set_err("failed to open " + path); return VOSS_ERR_INTERNAL; }
    std::fprintf(f, "pi_4_1,pi_4_3,difference\n%llu,%llu,%lld\n", (unsigned long long)c1, (unsigned long long)c3,
                 (long long)c3 - (long long)c1);
    std::fclose(f);
  }
  uint64_t n = 0;
  rc = voss_primes_ctx_large_gaps_count(ctx, &n);
  if (rc != VOSS_OK) return rc;
  {
    const std::string path = p + "_large_gaps.csv";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) { set_err("failed to open " + path); return VOSS_ERR_INTERNAL; }
    std::fprintf(f, "position,gap\n");
    for (uint64_t i = 0; i < n; ++i)
      std::fprintf(f, "%llu,%u\n", (unsigned long long)ctx->large[i].position, (unsigned)ctx->large[i].gap);
    std::fclose(f);
  }
  voss_primes_stats_t s;
  rc = voss_primes_ctx_statistics(ctx, &s);
  if (rc != VOSS_OK) return rc;
  {
    const std::string path = p + "_stats.csv";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) { set_err("failed to open " + path); return VOSS_ERR_INTERNAL; }
    std::fprintf(f, "metric,value\nmean_gap,%.6f\nstd_dev,%.6f\nskewness,%.6f\nkurtosis,%.6f\ntotal_gaps,%llu\n",
                 s.mean_gap, s.std_dev, s.skewness, s.kurtosis, (unsigned long long)s.total_gaps);
    std::fclose(f);
  }
  return VOSS_OK;
}

// ====================================================================== ranges
// range.cu:99-262: a < 2 becomes 2; b <= 100 by trial division (2, 3, 5
// included); otherwise the wheel bitmap, which has no 2, 3, 5 — the reference
// omits them for b > 100 and this keeps that behaviour.
extern "C" int voss_primes_in_range_with_limit(uint64_t a, uint64_t b, uint64_t max_count, uint64_t** out_array,
                                               uint64_t* out_count) {
  if (!out_array || !out_count) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  *out_array = nullptr;
  *out_count = 0;
  if (b > kMaxN) { set_err("b exceeds maximum (10^14)"); return VOSS_ERR_OUT_OF_RANGE; }
  if (max_count == 0) { set_err("max_count must be > 0"); return VOSS_ERR_INVALID_ARG; }
  return guarded([&] {
    if (a < 2) a = 2;
    std::vector<uint64_t> res;
    if (b < a) return VOSS_OK;
    if (b <= 100) {
      for (uint64_t n = a; n <= b; ++n) {
        bool pr = n >= 2;
        for (uint64_t d = 2; d * d <= n; ++d)
          if (n % d == 0) { pr = false; break; }
        if (pr) res.push_back(n);
      }
    } else {
      Engine& e = eng();
      std::lock_guard<std::mutex> lk(e.mu);
      bool of = false;
      res = e.primes_in_range(std::max<uint64_t>(a, 7), b, max_count, &of);
      if (of) { set_err("Range contains more primes than max_count"); return VOSS_ERR_OUT_OF_MEMORY; }
    }
    if (res.size() > max_count) { set_err("Range contains more primes than max_count"); return VOSS_ERR_OUT_OF_MEMORY; }
    if (res.empty()) return VOSS_OK;
    *out_array = to_malloc(res);
    *out_count = res.size();
    return VOSS_OK;
  });
}

extern "C" int voss_primes_in_range(uint64_t a, uint64_t b, uint64_t** out_array, uint64_t* out_count) {
  return voss_primes_in_range_with_limit(a, b, 10000000ull, out_array, out_count);  // range.cu:267-271
}

// ====================================================================== nth / next / prev / is_prime
extern "C" int voss_primes_nth(uint64_t n, uint64_t* out_prime) {  // nth.cu:40-103
  if (!out_prime) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (n == 0) { set_err("n must be >= 1 (1-indexed)"); return VOSS_ERR_INVALID_N; }
  if (n > 50000000ull) { set_err("n exceeds maximum (50,000,000)"); return VOSS_ERR_OUT_OF_RANGE; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    *out_prime = e.nth_prime(n);
    return VOSS_OK;
  });
}

extern "C" int voss_primes_next(uint64_t x, uint64_t* out_prime) {  // next_prev.cu:22-49
  if (!out_prime) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  uint64_t r = 0;
  if (!voss::host_next_prime(x, &r)) { set_err("no prime found within search limit"); return VOSS_ERR_OUT_OF_RANGE; }
  *out_prime = r;
  return VOSS_OK;
}

extern "C" int voss_primes_prev(uint64_t x, uint64_t* out_prime) {  // next_prev.cu:51-83
  if (!out_prime) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (x <= 2) { set_err("no prime < " + std::to_string(x)); return VOSS_ERR_INVALID_N; }
  uint64_t r = 0;
  if (!voss::host_prev_prime(x, &r)) { set_err("no prime found within search limit"); return VOSS_ERR_OUT_OF_RANGE; }
  *out_prime = r;
  return VOSS_OK;
}

extern "C" int voss_primes_is_prime(uint64_t x, int* out) {  // is_prime.cpp:11-18
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  *out = voss::host_is_prime(x) ? 1 : 0;
  return VOSS_OK;
}

// ====================================================================== Goldbach
extern "C" int voss_primes_goldbach_count(uint64_t n, uint64_t* out) {  // goldbach.cpp:47-69
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (n < 4 || (n & 1)) { set_err("n must be even and >= 4"); return VOSS_ERR_INVALID_N; }
  if (n > 100000000ull) { set_err("n exceeds maximum (10^8, CPU-only for now)"); return VOSS_ERR_OUT_OF_RANGE; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    *out = e.goldbach_counts(&n, 1)[0];
    return VOSS_OK;
  }, VOSS_ERR_INTERNAL);
}

extern "C" int voss_primes_goldbach_partitions(uint64_t n, uint64_t** out_array, uint64_t* out_count) {
  if (!out_array || !out_count) { set_err("out pointers are null"); return VOSS_ERR_INVALID_ARG; }  // goldbach.cpp:71-114
  *out_array = nullptr;
  *out_count = 0;
  if (n < 4 || (n & 1)) { set_err("n must be even and >= 4"); return VOSS_ERR_INVALID_N; }
  if (n > 100000000ull) { set_err("n exceeds maximum (10^8, CPU-only for now)"); return VOSS_ERR_OUT_OF_RANGE; }
  return guarded([&] {
    std::vector<uint64_t> flat;
    if (n == 4) {
      flat = {2, 2};
    } else {
      std::vector<uint64_t> ps;
      {
        Engine& e = eng();
        std::lock_guard<std::mutex> lk(e.mu);
        bool of = false;
        ps = e.primes_in_range(3, n - 3, ~0ull, &of);
      }
      // two pointers over the ascending list: p1 ascending, p2 = n - p1 descending
      size_t j = ps.size();
      for (size_t i = 0; i < ps.size() && ps[i] <= n / 2; ++i) {
        const uint64_t want = n - ps[i];
        while (j > 0 && ps[j - 1] > want) --j;
        if (j > 0 && ps[j - 1] == want) { flat.push_back(ps[i]); flat.push_back(want); }
      }
    }
    if (flat.empty()) return VOSS_OK;
    *out_array = to_malloc(flat);
    *out_count = flat.size() / 2;
    return VOSS_OK;
  }, VOSS_ERR_INTERNAL);
}

// The wheel bitmap of [0, N] (N/30 bytes) must fit the device: N <= 1e12.
constexpr uint64_t kGoldbachMaxN = 1000000000000ull;

extern "C" int voss_primes_goldbach_test_upto(uint64_t N, uint64_t** out_ce, uint64_t* out_count,
                                              uint64_t max_ce) {  // goldbach_upto.cu:48-150
  if (!out_ce || !out_count) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  *out_ce = nullptr;
  *out_count = 0;
  if (N < 4) { set_err("N must be >= 4"); return VOSS_ERR_INVALID_N; }
  if (max_ce == 0) return VOSS_OK;
  if (N > kGoldbachMaxN) { set_err("N exceeds 10^12 (the prime bitmap of [0, N] must fit the device)"); return VOSS_ERR_OUT_OF_RANGE; }
  return guarded([&] {
    std::vector<uint64_t> ce;
    {
      Engine& e = eng();
      std::lock_guard<std::mutex> lk(e.mu);
      ce = e.goldbach_upto(N, max_ce);
    }
    if (ce.empty()) return VOSS_OK;
    *out_ce = to_malloc(ce);
    *out_count = ce.size();
    return VOSS_OK;
  }, VOSS_ERR_INTERNAL);
}

extern "C" int voss_amd_goldbach_test_upto_pmax(uint64_t N, uint64_t pmax, uint64_t scan_limit, uint64_t max_ce,
                                                uint64_t** out_ce, uint64_t* out_count, uint64_t* out_total) {
  if (!out_ce || !out_count) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  *out_ce = nullptr;
  *out_count = 0;
  if (N < 4) { set_err("N must be >= 4"); return VOSS_ERR_INVALID_N; }
  if (N > kGoldbachMaxN) { set_err("N exceeds 10^12 (the prime bitmap of [0, N] must fit the device)"); return VOSS_ERR_OUT_OF_RANGE; }
  return guarded([&] {
    std::vector<uint64_t> ce;
    uint64_t total = 0;
    {
      Engine& e = eng();
      std::lock_guard<std::mutex> lk(e.mu);
      ce = e.goldbach_upto(N, max_ce, pmax, scan_limit, out_total ? &total : nullptr);
    }
    if (out_total) *out_total = total;
    if (ce.empty()) return VOSS_OK;
    *out_ce = to_malloc(ce);
    *out_count = ce.size();
    return VOSS_OK;
  }, VOSS_ERR_INTERNAL);
}

// ====================================================================== factorize / Mersenne / Fermat
extern "C" int voss_primes_factorize(uint64_t x, uint64_t** out_array, uint64_t* out_count) {
  if (!out_array || !out_count) { set_err("out pointers are null"); return VOSS_ERR_INVALID_ARG; }
  *out_array = nullptr;
  *out_count = 0;
  if (x < 2) { set_err("x must be >= 2"); return VOSS_ERR_INVALID_N; }
  return guarded([&] {
    std::vector<uint64_t> f = voss::host_factorize(x);
    *out_array = to_malloc(f);
    *out_count = f.size();
    return VOSS_OK;
  }, VOSS_ERR_INTERNAL);
}

extern "C" int voss_primes_is_mersenne_prime(uint32_t p, int* out) {
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (p < 2) { set_err("p must be >= 2"); return VOSS_ERR_INVALID_N; }  // mersenne_fermat.cpp
  if (p > 63) { set_err("p must be <= 63 (uint64 limit)"); return VOSS_ERR_OUT_OF_RANGE; }
  *out = voss::host_lucas_lehmer(p) ? 1 : 0;
  return VOSS_OK;
}

extern "C" int voss_primes_is_fermat_prime(uint32_t n, int* out) {
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (n > 5) { set_err("n must be <= 5 (uint64 limit for 2^(2^n)+1)"); return VOSS_ERR_OUT_OF_RANGE; }
  *out = voss::host_is_prime((1ull << (1u << n)) + 1) ? 1 : 0;
  return VOSS_OK;
}

// ====================================================================== extensions (voss_amd.h)
extern "C" int voss_amd_full_stats(uint64_t N, uint64_t cp_interval, uint32_t large_threshold,
                                   voss_amd_full_stats_t* out, uint64_t* hist, uint64_t* checkpoints,
                                   uint64_t cp_cap, voss_amd_gap_t* large, uint64_t large_cap,
                                   voss_amd_gap_t* records, uint64_t rec_cap) {
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (N < 2) { set_err("N must be >= 2"); return VOSS_ERR_INVALID_N; }
  if (N > kMaxN) { set_err("N exceeds maximum (10^14)"); return VOSS_ERR_OUT_OF_RANGE; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    voss::SweepStats S = e.sweep_stats(N, cp_interval, large_threshold, voss::ProgressFn());
    out->prime_count = S.prime_count;
    out->pi_4_1 = S.pi_4_1;
    out->pi_4_3 = S.pi_4_3;
    out->pi_3_1 = S.pi_3_1;
    out->pi_3_2 = S.pi_3_2;
    out->last_prime = S.last_prime;
    out->max_gap = S.max_gap;
    out->n_large = S.large_total;
    out->n_records = S.records.size();
    out->n_checkpoints = S.checkpoints.size() / 5;
    if (hist) {
      const size_t n = std::min<size_t>(kMaxGap, S.hist.size());
      std::memcpy(hist, S.hist.data(), n * sizeof(uint64_t));
      std::memset(hist + n, 0, (kMaxGap - n) * sizeof(uint64_t));
    }
    if (checkpoints)
      std::memcpy(checkpoints, S.checkpoints.data(),
                  std::min<uint64_t>(cp_cap, S.checkpoints.size() / 5) * 5 * sizeof(uint64_t));
    if (large)
      for (uint64_t i = 0; i < std::min<uint64_t>(large_cap, S.large.size()); ++i)
        large[i] = {S.large[i].position, S.large[i].gap, 0};
    if (records)
      for (uint64_t i = 0; i < std::min<uint64_t>(rec_cap, S.records.size()); ++i)
        records[i] = {S.records[i].position, S.records[i].gap, 0};
    return VOSS_OK;
  });
}

// batch results are written by the engine in place: the two layouts agree
static_assert(sizeof(voss::RangeOut) == sizeof(voss_amd_range_t), "RangeOut layout");
static_assert(offsetof(voss::RangeOut, count) == offsetof(voss_amd_range_t, count) &&
                  offsetof(voss::RangeOut, first) == offsetof(voss_amd_range_t, first) &&
                  offsetof(voss::RangeOut, last) == offsetof(voss_amd_range_t, last) &&
                  offsetof(voss::RangeOut, hist) == offsetof(voss_amd_range_t, hist) &&
                  offsetof(voss::RangeOut, pad) == offsetof(voss_amd_range_t, pad),
              "RangeOut layout");

extern "C" int voss_amd_batch_ranges(const uint64_t* a, const uint64_t* b, uint64_t n, voss_amd_range_t* out) {
  if ((!a || !b || !out) && n) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  for (uint64_t i = 0; i < n; ++i)
    if (b[i] > kMaxN) { set_err("b exceeds maximum (10^14)"); return VOSS_ERR_OUT_OF_RANGE; }
  return guarded([&] {
    // the engine writes each record once, straight into out (same layout)
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    e.batch_ranges(a, b, n, reinterpret_cast<voss::RangeOut*>(out));
    return VOSS_OK;
  });
}

extern "C" int voss_amd_range_stats(uint64_t a, uint64_t b, voss_amd_range_t* out) {
  return voss_amd_batch_ranges(&a, &b, 1, out);
}

extern "C" int voss_amd_count_range(uint64_t a, uint64_t b, uint64_t* out) {
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (b > kMaxN) { set_err("b exceeds maximum (10^14)"); return VOSS_ERR_OUT_OF_RANGE; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    *out = e.count_range(a, b);
    return VOSS_OK;
  });
}

extern "C" int voss_amd_count_range_noalloc(uint64_t a, uint64_t b, uint64_t* d_out, void* stream) {
  if (!d_out) { set_err("d_out is null"); return VOSS_ERR_INVALID_ARG; }
  if (b > kMaxN) { set_err("b exceeds maximum (10^14)"); return VOSS_ERR_OUT_OF_RANGE; }
  return guarded([&] {
    Engine* e = Engine::peek();  // never initialises (that allocates)
    if (!e) {
      set_err("the library is not initialised on this device: call a device entry point (voss_amd_device_info) first");
      return VOSS_ERR_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lk(e->mu);
    e->count_range_noalloc(a, b, reinterpret_cast<unsigned long long*>(d_out), static_cast<hipStream_t>(stream));
    return VOSS_OK;
  });
}

extern "C" int voss_amd_lds_probe(int kind, uint32_t iters, double* ops_per_clk_cu, double* clock_hz, double* ms) {
  if (!ops_per_clk_cu || !clock_hz || !ms) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    const Engine::LdsProbe r = e.lds_probe(kind, iters);
    *ops_per_clk_cu = r.ops_per_clk_cu;
    *clock_hz = r.clock_hz;
    *ms = r.ms;
    return VOSS_OK;
  });
}

extern "C" int voss_amd_is_prime_batch(const uint64_t* x, uint64_t n, uint8_t* out) {
  if ((!x || !out) && n) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    e.is_prime_batch(x, n, out);
    return VOSS_OK;
  });
}

extern "C" int voss_amd_next_prime_batch(const uint64_t* x, uint64_t n, uint64_t* out) {
  if ((!x || !out) && n) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    e.next_prev_batch(x, n, out, +1);
    return VOSS_OK;
  });
}

extern "C" int voss_amd_prev_prime_batch(const uint64_t* x, uint64_t n, uint64_t* out) {
  if ((!x || !out) && n) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    e.next_prev_batch(x, n, out, -1);
    return VOSS_OK;
  });
}

extern "C" int voss_amd_nth_prime(uint64_t n, uint64_t* out) {
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (n == 0) { set_err("n must be >= 1 (1-indexed)"); return VOSS_ERR_INVALID_N; }
  if (n > 1000000000000ull) { set_err("n exceeds maximum (10^12)"); return VOSS_ERR_OUT_OF_RANGE; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    *out = e.nth_prime(n);
    return VOSS_OK;
  });
}

extern "C" int voss_amd_goldbach_counts(const uint64_t* ns, uint64_t k, uint64_t* out) {
  if ((!ns || !out) && k) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  for (uint64_t i = 0; i < k; ++i) {
    if (ns[i] < 4 || (ns[i] & 1)) { set_err("n must be even and >= 4"); return VOSS_ERR_INVALID_N; }
    if (ns[i] > 10000000000ull) { set_err("n exceeds maximum (10^10)"); return VOSS_ERR_OUT_OF_RANGE; }
  }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    std::vector<uint64_t> r = e.goldbach_counts(ns, k);
    std::memcpy(out, r.data(), k * 8);
    return VOSS_OK;
  });
}

extern "C" int voss_amd_set_config_for(int kind, uint32_t seg_bytes, int block, int wgs_per_cu, uint32_t tier1_p) {
  if (kind < -1 || kind > 1) { set_err("kind must be -1 (both), 0 (sweep) or 1 (batch)"); return VOSS_ERR_INVALID_ARG; }
  return guarded([&] {
    Engine& e = eng();
    std::lock_guard<std::mutex> lk(e.mu);
    voss::TuneConfig c;
    c.seg_bytes = seg_bytes;
    c.block = block;
    c.wgs_per_cu = wgs_per_cu;
    c.tier1_p = tier1_p;
    e.set_config(c, kind);
    return VOSS_OK;
  });
}

extern "C" int voss_amd_set_config(uint32_t seg_bytes, int block, int wgs_per_cu, uint32_t tier1_p) {
  return voss_amd_set_config_for(-1, seg_bytes, block, wgs_per_cu, tier1_p);
}

extern "C" int voss_amd_get_config_for(int kind, uint32_t* seg_bytes, int* block, int* wgs_per_cu, uint32_t* tier1_p) {
  if (kind < 0 || kind > 1) { set_err("kind must be 0 (sweep) or 1 (batch)"); return VOSS_ERR_INVALID_ARG; }
  return guarded([&] {
    const voss::TuneConfig& c = eng().config(kind);
    if (seg_bytes) *seg_bytes = c.seg_bytes;
    if (block) *block = c.block;
    if (wgs_per_cu) *wgs_per_cu = c.wgs_per_cu;
    if (tier1_p) *tier1_p = c.tier1_p;
    return VOSS_OK;
  });
}

extern "C" int voss_amd_get_config(uint32_t* seg_bytes, int* block, int* wgs_per_cu, uint32_t* tier1_p) {
  return guarded([&] {
    const voss::TuneConfig& c = eng().config(0);
    if (seg_bytes) *seg_bytes = c.seg_bytes;
    if (block) *block = c.block;
    if (wgs_per_cu) *wgs_per_cu = c.wgs_per_cu;
    if (tier1_p) *tier1_p = c.tier1_p;
    return VOSS_OK;
  });
}

extern "C" int voss_amd_last_kernel_ms(double* sieve_ms, double* stitch_ms, double* total_ms, uint64_t* segments) {
  return guarded([&] {
    const voss::KernelTimes t = eng().last_times();
    if (sieve_ms) *sieve_ms = t.sieve_ms;
    if (stitch_ms) *stitch_ms = t.stitch_ms + t.rescan_ms;
    if (total_ms) *total_ms = t.total_ms;
    if (segments) *segments = t.segments;
    return VOSS_OK;
  });
}

extern "C" int voss_amd_device_info(int* num_cus, char* arch, int arch_len) {
  return guarded([&] {
    Engine& e = eng();
    if (num_cus) *num_cus = e.num_cus();
    if (arch && arch_len > 0) {
      std::strncpy(arch, e.device_name().c_str(), (size_t)arch_len - 1);
      arch[arch_len - 1] = 0;
    }
    return VOSS_OK;
  });
}

extern "C" int voss_amd_synchronize(void) {
  return guarded([&] {
    voss::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return VOSS_OK;
  });
}

// ====================================================================== serving
namespace {
std::mutex g_srv_mu;
// shared so that a wait on one thread keeps the server alive while another
// thread stops or replaces it (its scheduler drains every queued request
// before the last reference lets it go)
std::shared_ptr<voss::Server> g_srv;

voss::ServeConfig to_serve_config(const voss_amd_server_config_t* c) {
  voss::ServeConfig s;
  if (!c) return s;
  s.slots = c->slots;
  s.max_batch = c->max_batch;
  s.hbm_reserve = c->hbm_reserve;
  s.hbm_budget = c->hbm_budget;
  s.nth_max = c->nth_max;
  s.goldbach_max = c->goldbach_max;
  s.plain = c->plain != 0;
  return s;
}

void to_request(const voss_amd_request_t& in, voss::Request& r) {
  r.kind = in.kind;
  r.nx = in.n;
  r.a = in.a;
  r.b = in.b;
  std::memcpy(r.x, in.x, sizeof(r.x));
}

void to_response(const voss::Response& r, voss_amd_response_t* o) {
  std::memset(o, 0, sizeof(*o));
  o->status = r.status;
  o->value = r.value;
  o->first = r.first;
  o->last = r.last;
  std::memcpy(o->hist, r.hist, sizeof(o->hist));
  o->latency_ms = r.latency_ms;
  o->tick = r.batch;
}
}  // namespace

extern "C" int voss_amd_server_default_config(voss_amd_server_config_t* cfg) {
  if (!cfg) { set_err("cfg pointer is null"); return VOSS_ERR_INVALID_ARG; }
  const voss::ServeConfig d;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->slots = d.slots;
  cfg->max_batch = d.max_batch;
  cfg->hbm_reserve = d.hbm_reserve;
  cfg->hbm_budget = d.hbm_budget;
  cfg->nth_max = d.nth_max;
  cfg->goldbach_max = d.goldbach_max;
  cfg->plain = d.plain;
  return VOSS_OK;
}

extern "C" int voss_amd_server_start(const voss_amd_server_config_t* cfg) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_srv_mu);
    g_srv.reset();
    g_srv = std::make_shared<voss::Server>(to_serve_config(cfg));
    return VOSS_OK;
  });
}

extern "C" int voss_amd_server_stop(void) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_srv_mu);
    g_srv.reset();
    return VOSS_OK;
  });
}

extern "C" int voss_amd_server_submit(const voss_amd_request_t* req, uint64_t* ticket) {
  if (!req || !ticket) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_srv_mu);
    if (!g_srv) throw Error(VOSS_ERR_INVALID_ARG, "server not started");
    voss::Request r;
    to_request(*req, r);
    *ticket = g_srv->submit(r, voss::Server::Clock::now());
    return VOSS_OK;
  });
}

extern "C" int voss_amd_server_wait(uint64_t ticket, voss_amd_response_t* out) {
  if (!out) { set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  return guarded([&] {
    std::shared_ptr<voss::Server> s;
    {
      std::lock_guard<std::mutex> lk(g_srv_mu);
      s = g_srv;
    }
    if (!s) throw Error(VOSS_ERR_INVALID_ARG, "server not started");
    to_response(s->wait(ticket), out);
    return VOSS_OK;
  });
}

extern "C" int voss_amd_server_replay(const voss_amd_request_t* reqs, const double* arrival_s, uint64_t n,
                                      voss_amd_response_t* out, voss_amd_serve_stats_t* stats) {
  if ((!reqs || !arrival_s || !out) && n) { set_err("null pointer"); return VOSS_ERR_INVALID_ARG; }
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_srv_mu);
    if (!g_srv) throw Error(VOSS_ERR_INVALID_ARG, "server not started");
    voss::Server& S = *g_srv;
    using Clock = voss::Server::Clock;
    const voss::ServeStats before = S.stats();
    std::vector<uint64_t> tickets(n);
    std::vector<Clock::time_point> when(n);
    const Clock::time_point t0 = Clock::now() + std::chrono::milliseconds(2);
    for (uint64_t i = 0; i < n; ++i) {
      when[i] = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
      // coarse sleep, then spin the last 200 us: arrivals within ~10 us of schedule
      const auto early = when[i] - std::chrono::microseconds(200);
      if (Clock::now() < early) std::this_thread::sleep_until(early);
      while (Clock::now() < when[i]) {
      }
      voss::Request r;
      to_request(reqs[i], r);
      tickets[i] = S.submit(r, when[i]);
    }
    std::vector<double> lat(n);
    double end_s = 0;
    for (uint64_t i = 0; i < n; ++i) {
      voss::Response r = S.wait(tickets[i]);
      to_response(r, out + i);
      lat[i] = r.latency_ms;
      end_s = std::max(end_s, std::chrono::duration<double>(when[i] - t0).count() + r.latency_ms / 1e3);
    }
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      const voss::ServeStats after = S.stats();
      stats->wall_s = end_s - (n ? std::chrono::duration<double>(when[0] - t0).count() : 0.0);
      stats->qps = stats->wall_s > 0 ? (double)n / stats->wall_s : 0;
      if (n) {
        std::vector<double> sl = lat;
        std::sort(sl.begin(), sl.end());
        auto pct = [&](double q) {  // nearest rank
          size_t k = (size_t)std::ceil(q * (double)n);
          if (k < 1) k = 1;
          return sl[k - 1];
        };
        stats->p50_ms = pct(0.50);
        stats->p99_ms = pct(0.99);
        stats->max_ms = sl.back();
        double sum = 0;
        for (double v : sl) sum += v;
        stats->mean_ms = sum / (double)n;
      }
      stats->warmup_s = after.warmup_s;
      stats->ticks = after.ticks - before.ticks;
      stats->max_tick = after.max_tick;
      stats->deferred_by_memory = after.deferred_by_memory - before.deferred_by_memory;
      stats->peak_inflight_bytes = after.peak_inflight_bytes;
    }
    return VOSS_OK;
  });
}
