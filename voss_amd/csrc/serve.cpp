// serve.cpp — the continuous-batching request server (declared in serve.h).
//
// The reference serves every query as its own call (library/src/primes/
// prime_count.cu:296-320, in_range.cpp, nth.cu:40-103, next_prev.cu,
// goldbach.cpp:21-69, core/is_prime.cpp); this server keeps the state those
// calls rebuild each time resident instead:
//   - the sieving-prime table (Engine, built once at start-up);
//   - a prefix-count index of sweep segments up to the nth bound, so nth(n)
//     is a binary search on the host plus one segment sieve + select;
//   - the wheel bitmap of [0, goldbach_max] for Goldbach partition counts.
// Results are the ones the plain entry points return for the same request.
#include "serve.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

#include "kernels.h"
#include "misc_kernels.h"
#include "voss/voss.h"

namespace voss {

namespace {
constexpr uint64_t kMaxRangeB = 100000000000000ull;  // 1e14, as the range entry points
constexpr size_t kInPerReq = 8 * kReqMaxX;           // staging bytes per request, worst kind (is_prime)
constexpr size_t kOutPerReq = kReqMaxX;

uint32_t pow2_ge(uint64_t v, uint32_t lo) {
  uint64_t c = lo;
  while (c < v) c <<= 1;
  return (uint32_t)c;
}

uint64_t rosser_bound(uint64_t n) {  // p_n < n (ln n + ln ln n) for n >= 6
  const double nf = (double)std::max<uint64_t>(n, 6);
  return (uint64_t)(nf * (std::log(nf) + std::log(std::log(nf)))) + 64;
}
}  // namespace

struct Server::Slot {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  uint64_t tick = 0, bytes = 0;
  std::vector<Item> items;
  std::vector<uint32_t> by_kind[kReqKinds];
  size_t in_off[kReqKinds] = {0}, out_off[kReqKinds] = {0};
  size_t in_cap = 0, out_cap = 0;
  void* h_in = nullptr;
  void* h_out = nullptr;
  DevBuf d_in, d_out;
  std::vector<std::unique_ptr<BatchPlan>> plans;
  BatchPlan* plan = nullptr;
  Engine::BatchLayout layout;
  std::vector<uint64_t> ra, rb;
  std::vector<RangeOut> rout;
  std::vector<uint32_t> nth_dev;  // positions (within by_kind[kReqNth]) computed on the device
  ~Slot() {
    plans.clear();
    if (h_in) (void)hipHostFree(h_in);
    if (h_out) (void)hipHostFree(h_out);
    if (done) (void)hipEventDestroy(done);
    if (st) (void)hipStreamDestroy(st);
  }
};

Server::Server(const ServeConfig& c) : cfg_(c), eng_(Engine::get()) {
  static std::atomic<uint64_t> generations{0};
  gen_ = (++generations) & 0xFFFFFFull;
  if (cfg_.plain) {
    cfg_.slots = 1;
    cfg_.max_batch = 1;
  }
  if (cfg_.slots < 1 || cfg_.slots > 8) throw Error(VOSS_ERR_INVALID_ARG, "slots must be 1..8");
  if (cfg_.max_batch < 1 || cfg_.max_batch > 65535) throw Error(VOSS_ERR_INVALID_ARG, "max_batch must be 1..65535");
  if (cfg_.nth_max > 1000000000000ull) throw Error(VOSS_ERR_OUT_OF_RANGE, "nth_max exceeds 10^12");
  if (cfg_.goldbach_max > 10000000000ull) throw Error(VOSS_ERR_OUT_OF_RANGE, "goldbach_max exceeds 10^10");
  const auto t0 = Clock::now();
  {
    std::lock_guard<std::mutex> lk(eng_.mu);
    rcfg_ = eng_.config(1);
    icfg_ = eng_.config(0);
    // nth index: per-segment counts of [7, rosser(nth_max)], prefix-summed
    // (built once per process and shared by later servers)
    if (cfg_.nth_max >= 4) {
      idx_ = eng_.nth_index(rosser_bound(cfg_.nth_max), icfg_);
      if (idx_->cum.back() < cfg_.nth_max) throw Error(VOSS_ERR_INTERNAL, "nth index does not reach nth_max");
    }
    if (cfg_.goldbach_max >= 4) {
      uint64_t kb = 0;
      eng_.bitmap_range(7, std::max<uint64_t>(cfg_.goldbach_max, 7), gb_bitmap_, &kb);
    }
  }
  for (int i = 0; i < cfg_.slots; ++i) {
    std::unique_ptr<Slot> s(new Slot());
    hip_check(hipStreamCreate(&s->st), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&s->done, hipEventDisableTiming), "hipEventCreate");
    s->in_cap = 64 + (size_t)cfg_.max_batch * kInPerReq;
    s->out_cap = (size_t)cfg_.max_batch * kOutPerReq + 64;
    hip_check(hipHostMalloc(&s->h_in, s->in_cap, hipHostMallocDefault), "hipHostMalloc");
    hip_check(hipHostMalloc(&s->h_out, s->out_cap, hipHostMallocDefault), "hipHostMalloc");
    s->d_in = DevBuf(s->in_cap);
    s->d_out = DevBuf(s->out_cap);
    slots_.push_back(std::move(s));
  }
  st_.warmup_s = std::chrono::duration<double>(Clock::now() - t0).count();
  thread_ = std::thread([this] { run(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  inbox_cv_.notify_all();
  if (thread_.joinable()) thread_.join();
  for (auto& s : slots_)
    if (s->st) (void)hipStreamSynchronize(s->st);
}

int Server::validate(const Request& r) const {
  switch (r.kind) {
    case kReqRange:
      return r.b > kMaxRangeB ? VOSS_ERR_OUT_OF_RANGE : VOSS_OK;
    case kReqNth:
      if (r.a == 0) return VOSS_ERR_INVALID_N;
      return r.a > cfg_.nth_max ? VOSS_ERR_OUT_OF_RANGE : VOSS_OK;
    case kReqNext:
    case kReqPrev:
      return VOSS_OK;
    case kReqIsPrime:
      return r.nx > kReqMaxX ? VOSS_ERR_INVALID_ARG : VOSS_OK;
    case kReqGoldbach:
      if (r.a < 4 || (r.a & 1)) return VOSS_ERR_INVALID_N;
      return r.a > cfg_.goldbach_max ? VOSS_ERR_OUT_OF_RANGE : VOSS_OK;
    default:
      return VOSS_ERR_INVALID_ARG;
  }
}

uint64_t Server::bytes_of(const Request& r) const {
  if (r.kind == kReqRange) {
    const uint64_t lo = std::max<uint64_t>(r.a, 7);
    const uint64_t S = rcfg_.seg_bytes;
    const uint64_t nseg = r.b >= lo ? ((r.b / 30 + 1) - lo / 30 + S - 1) / S : 0;
    return nseg * (sizeof(SegSummary) + 4) + sizeof(QueryDesc) + 24 + 4 * kQueryBins;
  }
  return kInPerReq + kOutPerReq;
}

uint64_t Server::submit(const Request& r, Clock::time_point arrival) {
  Item it;
  it.req = r;
  it.arrival = arrival;
  it.bytes = bytes_of(r);
  const int v = validate(r);
  std::unique_lock<std::mutex> lk(mu_);
  it.ticket = (gen_ << 40) | next_ticket_++;
  if (v != VOSS_OK) {
    Response& o = done_[it.ticket];
    o.status = v;
    o.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - arrival).count();
    lk.unlock();
    done_cv_.notify_all();
    return it.ticket;
  }
  inbox_.push_back(it);
  lk.unlock();
  inbox_cv_.notify_one();
  return it.ticket;
}

Response Server::wait(uint64_t ticket) {
  if ((ticket >> 40) != gen_) throw Error(VOSS_ERR_INVALID_ARG, "ticket was issued by another server");
  std::unique_lock<std::mutex> lk(mu_);
  done_cv_.wait(lk, [&] { return done_.count(ticket) != 0; });
  Response r = done_[ticket];
  done_.erase(ticket);
  return r;
}

ServeStats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  return st_;
}

void Server::complete(const Item& it, Response& r, uint64_t tick, Clock::time_point now) {
  r.latency_ms = std::chrono::duration<double, std::milli>(now - it.arrival).count();
  r.batch = tick;
  done_[it.ticket] = r;
}

// ------------------------------------------------------------------ one tick
void Server::launch(Slot& s) {
  for (auto& v : s.by_kind) v.clear();
  for (uint32_t i = 0; i < (uint32_t)s.items.size(); ++i) s.by_kind[s.items[i].req.kind].push_back(i);

  // ranges: one replay of a batched-sieve graph owned by this slot
  const auto& R = s.by_kind[kReqRange];
  s.plan = nullptr;
  s.layout = Engine::BatchLayout();
  s.ra.resize(R.size());
  s.rb.resize(R.size());
  for (size_t j = 0; j < R.size(); ++j) {
    s.ra[j] = s.items[R[j]].req.a;
    s.rb[j] = s.items[R[j]].req.b;
  }
  if (!R.empty()) {
    s.layout = eng_.layout_batch(s.ra.data(), s.rb.data(), R.size(), rcfg_.seg_bytes);
    if (s.layout.nq) {
      const uint32_t qcap = pow2_ge(s.layout.nq, 64), scap = pow2_ge(s.layout.nseg, 1024);
      for (auto& p : s.plans)
        if (p->qcap == qcap && p->scap == scap) s.plan = p.get();
      if (!s.plan) {
        if (s.plans.size() >= 6) s.plans.erase(s.plans.begin());
        // ranges inside the resident Goldbach bitmap read it instead of sieving
        const bool cache = cfg_.goldbach_max >= 4 && gb_bitmap_.p;
        s.plans.push_back(eng_.make_batch_plan(qcap, scap, rcfg_, s.st, cache ? gb_bitmap_.as<uint8_t>() : nullptr,
                                               cache ? cfg_.goldbach_max : 0));
        s.plan = s.plans.back().get();
      }
      Engine::stage_batch(*s.plan, s.layout, s.ra.data(), s.rb.data());
      hip_check(hipGraphLaunch(s.plan->exec, s.st), "hipGraphLaunch");
    }
  }

  // everything else shares one staging upload and one download
  uint8_t* hin = static_cast<uint8_t*>(s.h_in);
  uint32_t* hdr = reinterpret_cast<uint32_t*>(hin);
  std::memset(hdr, 0, 64);
  size_t in = 64, out = 0;
  auto region = [&](int kind, size_t in_bytes, size_t out_bytes) {
    s.in_off[kind] = in;
    s.out_off[kind] = out;
    in += (in_bytes + 7) & ~(size_t)7;
    out += (out_bytes + 7) & ~(size_t)7;
  };
  const size_t nN = s.by_kind[kReqNext].size(), nP = s.by_kind[kReqPrev].size();
  const size_t nI = s.by_kind[kReqIsPrime].size(), nG = s.by_kind[kReqGoldbach].size();
  // nth: n <= 3 answered on the host; the rest located by the index
  s.nth_dev.clear();
  for (uint32_t j = 0; j < (uint32_t)s.by_kind[kReqNth].size(); ++j)
    if (s.items[s.by_kind[kReqNth][j]].req.a > 3) s.nth_dev.push_back(j);
  const size_t nT = s.nth_dev.size();
  region(kReqNext, 8 * nN, 8 * nN);
  region(kReqPrev, 8 * nP, 8 * nP);
  region(kReqIsPrime, 8 * kReqMaxX * nI, kReqMaxX * nI);
  region(kReqGoldbach, 8 * nG, 8 * nG);
  region(kReqNth, 8 * nT, 8 * nT);
  if (in > s.in_cap || out > s.out_cap) throw Error(VOSS_ERR_INTERNAL, "tick exceeds its staging");
  auto put64 = [&](int kind, const std::vector<uint32_t>& idx, int which) {
    uint64_t* p = reinterpret_cast<uint64_t*>(hin + s.in_off[kind]);
    for (size_t j = 0; j < idx.size(); ++j) p[j] = which ? s.items[idx[j]].req.b : s.items[idx[j]].req.a;
  };
  put64(kReqNext, s.by_kind[kReqNext], 0);
  put64(kReqPrev, s.by_kind[kReqPrev], 0);
  put64(kReqGoldbach, s.by_kind[kReqGoldbach], 0);
  {
    uint64_t* p = reinterpret_cast<uint64_t*>(hin + s.in_off[kReqIsPrime]);
    for (size_t j = 0; j < nI; ++j) {
      const Request& r = s.items[s.by_kind[kReqIsPrime][j]].req;
      for (uint32_t t = 0; t < kReqMaxX; ++t) p[j * kReqMaxX + t] = t < r.nx ? r.x[t] : 0;
    }
  }
  {
    uint32_t* segs = reinterpret_cast<uint32_t*>(hin + s.in_off[kReqNth]);
    uint32_t* ranks = segs + nT;
    for (size_t j = 0; j < nT; ++j) {
      const uint64_t n = s.items[s.by_kind[kReqNth][s.nth_dev[j]]].req.a;
      // first segment whose prefix reaches n, minus one: cum[sg] < n <= cum[sg + 1]
      const std::vector<uint64_t>& cum = idx_->cum;
      const size_t sg = (size_t)(std::lower_bound(cum.begin(), cum.end(), n) - cum.begin()) - 1;
      segs[j] = (uint32_t)sg;
      ranks[j] = (uint32_t)(n - cum[sg]);
    }
    hdr[0] = (uint32_t)nT;
  }
  uint8_t* din = s.d_in.as<uint8_t>();
  uint8_t* dout = s.d_out.as<uint8_t>();
  if (in > 64) {
    hip_check(hipMemcpyAsync(din, hin, in, hipMemcpyHostToDevice, s.st), "H2D");
    if (nN)
      hip_check(launch_next_prev(reinterpret_cast<const uint64_t*>(din + s.in_off[kReqNext]), nN,
                                 reinterpret_cast<uint64_t*>(dout + s.out_off[kReqNext]), 1, s.st),
                "launch(next)");
    if (nP)
      hip_check(launch_next_prev(reinterpret_cast<const uint64_t*>(din + s.in_off[kReqPrev]), nP,
                                 reinterpret_cast<uint64_t*>(dout + s.out_off[kReqPrev]), -1, s.st),
                "launch(prev)");
    if (nI)
      hip_check(launch_is_prime(reinterpret_cast<const uint64_t*>(din + s.in_off[kReqIsPrime]), kReqMaxX * nI,
                                dout + s.out_off[kReqIsPrime], s.st),
                "launch(is_prime)");
    if (nG) {
      uint64_t gmax = 4;
      for (uint32_t i : s.by_kind[kReqGoldbach]) gmax = std::max(gmax, s.items[i].req.a);
      hip_check(launch_goldbach_count(gb_bitmap_.as<uint8_t>(),
                                      reinterpret_cast<const uint64_t*>(din + s.in_off[kReqGoldbach]), nG, gmax,
                                      ~0ull, reinterpret_cast<unsigned long long*>(dout + s.out_off[kReqGoldbach]), s.st),
                "launch(goldbach)");
    }
    if (nT) {
      const uint32_t nseg = (uint32_t)(idx_->cum.size() - 1);
      SieveArgs A = eng_.base_args(nseg, icfg_);
      A.seg_bytes = idx_->seg_bytes;
      A.k_base = 0;
      A.k_end = idx_->hi / 30 + 1;
      A.lo_val = 7;
      A.hi_val = idx_->hi;
      const uint32_t* segs = reinterpret_cast<const uint32_t*>(din + s.in_off[kReqNth]);
      hip_check(launch_select(icfg_.block, A, segs, segs + nT,
                              reinterpret_cast<unsigned long long*>(dout + s.out_off[kReqNth]),
                              reinterpret_cast<const uint32_t*>(din), (uint32_t)nT, s.st),
                "launch(select)");
    }
    if (out) hip_check(hipMemcpyAsync(s.h_out, dout, out, hipMemcpyDeviceToHost, s.st), "D2H");
  }
  hip_check(hipEventRecord(s.done, s.st), "hipEventRecord");
  s.busy = true;
}

void Server::finish(Slot& s) {
  const auto now = Clock::now();
  std::vector<Response> res(s.items.size());
  const auto& R = s.by_kind[kReqRange];
  if (!R.empty()) {
    s.rout.resize(R.size());
    Engine::read_batch(s.plan, s.layout, s.ra.data(), s.rb.data(), R.size(), s.rout.data());
    for (size_t j = 0; j < R.size(); ++j) {
      Response& o = res[R[j]];
      o.value = s.rout[j].count;
      o.first = s.rout[j].first;
      o.last = s.rout[j].last;
      std::memcpy(o.hist, s.rout[j].hist, sizeof(o.hist));
    }
  }
  const uint8_t* hout = static_cast<const uint8_t*>(s.h_out);
  auto get64 = [&](int kind, size_t j) { return reinterpret_cast<const uint64_t*>(hout + s.out_off[kind])[j]; };
  for (size_t j = 0; j < s.by_kind[kReqNext].size(); ++j) {
    Response& o = res[s.by_kind[kReqNext][j]];
    o.value = get64(kReqNext, j);
    if (!o.value) o.status = VOSS_ERR_OUT_OF_RANGE;  // no prime above x in uint64
  }
  for (size_t j = 0; j < s.by_kind[kReqPrev].size(); ++j) res[s.by_kind[kReqPrev][j]].value = get64(kReqPrev, j);
  for (size_t j = 0; j < s.by_kind[kReqGoldbach].size(); ++j)
    res[s.by_kind[kReqGoldbach][j]].value = get64(kReqGoldbach, j);
  for (size_t j = 0; j < s.by_kind[kReqIsPrime].size(); ++j) {
    const Request& r = s.items[s.by_kind[kReqIsPrime][j]].req;
    const uint8_t* b = hout + s.out_off[kReqIsPrime] + j * kReqMaxX;
    uint64_t m = 0;
    for (uint32_t t = 0; t < r.nx; ++t) m |= (uint64_t)(b[t] & 1) << t;
    res[s.by_kind[kReqIsPrime][j]].value = m;
  }
  {
    static const uint64_t small[4] = {0, 2, 3, 5};
    const auto& T = s.by_kind[kReqNth];
    for (size_t j = 0; j < T.size(); ++j) {
      const uint64_t n = s.items[T[j]].req.a;
      if (n <= 3) res[T[j]].value = small[n];
    }
    for (size_t j = 0; j < s.nth_dev.size(); ++j) {
      Response& o = res[T[s.nth_dev[j]]];
      o.value = get64(kReqNth, j);
      if (!o.value) o.status = VOSS_ERR_INTERNAL;  // the index and the segment disagree
    }
  }
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (size_t i = 0; i < s.items.size(); ++i) complete(s.items[i], res[i], s.tick, now);
  }
  done_cv_.notify_all();
  inflight_bytes_ -= s.bytes;
  s.items.clear();
  s.bytes = 0;
  s.busy = false;
}

void Server::fail_slot(Slot& s, int code) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    const auto now = Clock::now();
    for (auto& it : s.items) {
      Response r;
      r.status = code;
      complete(it, r, s.tick, now);
    }
  }
  done_cv_.notify_all();
  inflight_bytes_ -= s.bytes;
  s.items.clear();
  s.bytes = 0;
  s.busy = false;
}

// ------------------------------------------------------------------ scheduler
void Server::run() {
  size_t free_hbm = 0, total_hbm = 0;
  auto last_meminfo = Clock::now() - std::chrono::seconds(1);
  for (;;) {
    bool any_busy = false;
    for (auto& sp : slots_) any_busy |= sp->busy;
    {
      std::unique_lock<std::mutex> lk(mu_);
      if (!any_busy && pending_.empty() && inbox_.empty()) {
        if (stop_) break;
        inbox_cv_.wait_for(lk, std::chrono::milliseconds(2));
      }
      while (!inbox_.empty()) {
        pending_.push_back(inbox_.front());
        inbox_.pop_front();
      }
    }
    bool progressed = false;
    // completions
    for (auto& sp : slots_) {
      Slot& s = *sp;
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      if (q == hipSuccess) {
        try {
          finish(s);
        } catch (const std::exception&) {
          fail_slot(s, VOSS_ERR_INTERNAL);
        }
      } else {
        fail_slot(s, VOSS_ERR_CUDA);
      }
      if (!s.busy) {
        progressed = true;
        done_cv_.notify_all();
      }
    }
    // admissions: fill every free slot from the pending queue
    for (auto& sp : slots_) {
      Slot& s = *sp;
      if (s.busy || pending_.empty()) continue;
      const auto now = Clock::now();
      if (now - last_meminfo > std::chrono::milliseconds(1)) {
        if (hipMemGetInfo(&free_hbm, &total_hbm) != hipSuccess) free_hbm = 0;
        last_meminfo = now;
      }
      const uint64_t dev_room = free_hbm > cfg_.hbm_reserve ? free_hbm - cfg_.hbm_reserve : 0;
      const uint64_t budget = std::min<uint64_t>(cfg_.hbm_budget, dev_room);
      uint64_t bytes = 0;
      while (!pending_.empty() && s.items.size() < cfg_.max_batch) {
        Item& it = pending_.front();
        if (inflight_bytes_ + bytes + it.bytes > budget) {
          if (s.items.empty() && inflight_bytes_ == 0) {
            // can never fit: fail it rather than wait forever
            Response r;
            r.status = VOSS_ERR_OUT_OF_MEMORY;
            {
              std::lock_guard<std::mutex> lk(mu_);
              complete(it, r, 0, Clock::now());
            }
            done_cv_.notify_all();
            pending_.pop_front();
            continue;
          }
          if (!it.deferred) {  // count each request once, however long it waits
            it.deferred = true;
            std::lock_guard<std::mutex> lk(mu_);
            st_.deferred_by_memory++;
          }
          break;
        }
        bytes += it.bytes;
        s.items.push_back(it);
        pending_.pop_front();
      }
      if (s.items.empty()) continue;
      s.bytes = bytes;
      inflight_bytes_ += bytes;
      s.tick = ++tick_serial_;
      {
        std::lock_guard<std::mutex> lk(mu_);
        st_.ticks++;
        st_.requests += s.items.size();
        st_.max_tick = std::max<uint64_t>(st_.max_tick, s.items.size());
        st_.peak_inflight_bytes = std::max(st_.peak_inflight_bytes, inflight_bytes_);
      }
      try {
        launch(s);
      } catch (const Error& e) {
        (void)hipStreamSynchronize(s.st);
        fail_slot(s, e.code);
      } catch (const std::exception&) {
        (void)hipStreamSynchronize(s.st);
        fail_slot(s, VOSS_ERR_INTERNAL);
      }
      progressed = true;
    }
    if (!progressed) std::this_thread::yield();
  }
}

}  // namespace voss
