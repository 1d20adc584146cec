// serve.h — continuous-batching request server over the sieve engine.
//
// One scheduler thread owns `slots` HIP streams.  Requests submitted from any
// thread wait in an inbox; whenever a slot is free the scheduler admits the
// waiting requests in arrival order, as many as the slot's batch limit and
// the device-memory budget allow, and launches them on that slot's stream
// (ranges as one graph replay of the batched sieve, is_prime/next/prev as one
// Miller-Rabin launch each, Goldbach counts over a resident bitmap, nth as a
// select launch over a resident prefix-count index).  Requests that arrive
// while other slots are busy are admitted into the next free slot, so their
// work overlaps on the device.  A completed slot's results are written back
// and its waiters woken.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"

namespace voss {

enum ReqKind : uint32_t {
  kReqRange = 0,     // count + gap histogram + first/last over [a, b]
  kReqNth = 1,       // a-th prime (1-based)
  kReqNext = 2,      // smallest prime > a
  kReqPrev = 3,      // largest prime < a (0 when a <= 2)
  kReqIsPrime = 4,   // primality of x[0..nx), nx <= 64: bit i of value
  kReqGoldbach = 5,  // partition count of even a (p1 <= p2)
  kReqKinds = 6
};

constexpr uint32_t kReqMaxX = 64;

struct Request {
  uint32_t kind = 0;
  uint32_t nx = 0;
  uint64_t a = 0, b = 0;
  uint64_t x[kReqMaxX] = {0};
};

struct Response {
  int status = 0;         // VOSS_* code
  uint64_t value = 0;     // range: count; nth/next/prev: prime; is_prime: bitmask; goldbach: count
  uint64_t first = 0, last = 0;
  uint32_t hist[kQueryBins] = {0};
  double latency_ms = 0;  // completion minus arrival
  uint64_t batch = 0;     // serial number of the tick that carried it
};

struct ServeConfig {
  int slots = 4;                              // concurrent streams
  uint32_t max_batch = 1024;                  // requests per tick
  uint64_t hbm_reserve = 1ull << 30;          // left free on the device
  uint64_t hbm_budget = 8ull << 30;           // working bytes in flight across slots
  uint64_t nth_max = 1000000000000ull;        // largest n served by nth (index size)
  uint64_t goldbach_max = 100000000ull;       // largest n served by goldbach (bitmap size)
  bool plain = false;                         // one request per tick, one slot, wait per request
};

struct ServeStats {
  double warmup_s = 0;
  uint64_t ticks = 0, requests = 0, max_tick = 0;
  uint64_t peak_inflight_bytes = 0;
  uint64_t deferred_by_memory = 0;  // requests held back at least once because the budget was full
};

class Server {
 public:
  using Clock = std::chrono::steady_clock;
  explicit Server(const ServeConfig& c);  // builds the resident index and bitmap
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Enqueue; arrival is the time latency is measured from.
  uint64_t submit(const Request& r, Clock::time_point arrival);
  // Block until ticket completes; copies its response and forgets it.
  // Throws for a ticket issued by another server.
  Response wait(uint64_t ticket);
  ServeStats stats();
  const ServeConfig& config() const { return cfg_; }

  // Validation shared with submit (VOSS_* code; 0 = acceptable).
  int validate(const Request& r) const;

 private:
  struct Item {
    uint64_t ticket;
    Request req;
    Clock::time_point arrival;
    uint64_t bytes;  // device working bytes while in flight
    bool deferred = false;  // already held back once by the memory budget
  };
  struct Slot;

  void run();
  uint64_t bytes_of(const Request& r) const;
  void launch(Slot& s);
  void finish(Slot& s);
  void fail_slot(Slot& s, int code);  // completes every item of s with code
  void complete(const Item& it, Response& r, uint64_t tick, Clock::time_point now);

  ServeConfig cfg_;
  Engine& eng_;
  TuneConfig rcfg_;  // batched-range configuration
  TuneConfig icfg_;  // configuration the nth index was counted with
  // resident state
  std::shared_ptr<const NthIndex> idx_;  // nth prefix index (shared with the engine's cache)
  DevBuf gb_bitmap_;           // wheel bitmap of [0, goldbach_max]
  std::vector<std::unique_ptr<Slot>> slots_;

  std::mutex mu_;
  std::condition_variable inbox_cv_, done_cv_;
  std::deque<Item> inbox_;
  std::unordered_map<uint64_t, Response> done_;
  uint64_t next_ticket_ = 1;
  uint64_t gen_ = 0;  // server generation, in the ticket's top 24 bits
  bool stop_ = false;
  ServeStats st_;
  std::deque<Item> pending_;  // scheduler-owned
  uint64_t inflight_bytes_ = 0;
  uint64_t tick_serial_ = 0;
  std::thread thread_;
};

}  // namespace voss
